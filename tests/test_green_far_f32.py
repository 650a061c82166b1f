"""The FP32 far-field form of the mixed solve's Green's function
(green_far_f in src_amd/csrc/igfmath.hpp) against the FP64 form (green_far),
checked on the host: both are compiled with g++ from the header's own text and
evaluated over far points of octants with cell sides from 1e-9 to 1e3 m and
aspect ratios up to 1:50, where FP32 powers up to r^-9 would leave FP32's range
without the scaling to the largest cell side."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MAIN = r'''
#include <cstdio>
#include <cmath>
#define __host__
#define __device__
using std::fmax;
%s
int main() {
    const double hs[][3] = {{1e-5, 1.1e-5, 2e-5}, {1e-9, 1e-9, 5e-8}, {1e3, 7e2, 2e1}, {3e-4, 3e-4, 3e-4},
                            {2e-6, 1e-4, 3e-5}};
    double worst = 0.0;
    for (const auto &h : hs) {
        const GreenMoments m = green_moments(h[0], h[1], h[2]);
        const GreenMomentsF mf = green_moments_f(h[0], h[1], h[2]);
        for (int k = 0; k <= 256; k += 3)
            for (int j = 0; j <= 256; j += 5)
                for (int i = 0; i <= 256; i += 7) {
                    const double cx = i * h[0], cy = j * h[1], cz = k * h[2];
                    if (cx * cx + cy * cy + cz * cz < m.far2) continue;
                    const double g = green_far(m, cx, cy, cz);
                    const float f = green_far_f(mf, (float)(cx * mf.rhm), (float)(cy * mf.rhm), (float)(cz * mf.rhm));
                    if (!std::isfinite(f) || !(f > 0.0f)) { printf("bad %%g at %%d %%d %%d\n", (double)f, i, j, k); return 1; }
                    const double e = std::fabs((double)f - g) / g;
                    if (e > worst) worst = e;
                }
    }
    printf("%%.3e\n", worst);
    return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fp32_far_field_within_a_few_ulps_of_fp64(tmp_path):
    src = open(os.path.join(ROOT, "src_amd", "csrc", "igfmath.hpp")).read()
    i = src.index("constexpr double kGreenFarCells")
    j = src.index("// G at octant point (i, j, k) of the mixed-precision solve")
    path = tmp_path / "far.cpp"
    path.write_text(MAIN % src[i:j])
    exe = tmp_path / "far"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", str(exe), str(path)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    worst = float(out.stdout.strip())
    assert worst < 5e-7, worst  # a few FP32 ulps (2^-24 = 6e-8)
