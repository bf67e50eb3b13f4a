"""MI355X-native KinectFusion hot path (depth preprocessing, TSDF integration,
raycasting, ICP tracking) over libkinfu's HIP kernels for gfx950.

Submodules: ``kinfu`` (C-ABI binding and reference-named mirrors), ``synth``
(seeded synthetic streams and workload presets), ``metrics`` (ATE), ``build``.
"""
