"""The C++ drop-in header compiles into a reference-shaped driver and links
against libkinfu (no GPU call is made: the program only exercises argument
validation, which returns before touching the device)."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cstdio>
#include "kinfu.hpp"

int main() {
  kf_intrinsics k{525.f, 525.f, 319.5f, 239.5f, 640, 480};
  kf_icp_params bad{};  // levels = 0: rejected
  try {
    kinfu::ICP icp(k, bad);
    return 1;
  } catch (const kinfu::Error& e) {
    if (e.status != KF_ERR_INVALID_ARGUMENT) return 2;
  }
  kf_volume_desc d{};
  d.n = 12;  // not a multiple of 8
  kf_volume* v = nullptr;
  if (kf_volume_create(&d, &v) != KF_ERR_INVALID_ARGUMENT) return 3;
  kf_intrinsics k1 = kf_intrinsics_level(&k, 1);
  if (k1.width != 320 || k1.cx != 159.5f) return 4;
  std::printf("ok\n");
  return 0;
}
"""


def test_cpp_header_driver_compiles_and_links(kf):
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "drv.cpp")
        exe = os.path.join(td, "drv")
        open(src, "w").write(DRIVER)
        libdir = os.path.join(ROOT, "openfusion_amd")
        subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src, "-o", exe,
                        f"-L{libdir}", "-lkinfu", f"-Wl,-rpath,{libdir}"], check=True)
        out = subprocess.run([exe], capture_output=True, text=True)
        assert out.returncode == 0, (out.returncode, out.stderr)
        assert out.stdout.strip() == "ok"
