"""Properties of the tracker's re-sort key (cell_key_hilbert in
src_amd/csrc/common.hpp), checked on the host: the function is compiled with
g++ from the header's own text and walked over every cell of 2^b cubes."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MAIN = r'''
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <vector>
#define __host__
#define __device__
%s
int main() {
    for (int b = 1; b <= 6; ++b) {
        const int n = 1 << b;
        const size_t m = (size_t)n * n * n;
        std::vector<int> cx(m, -1), cy(m), cz(m);
        for (int x = 0; x < n; ++x)
            for (int y = 0; y < n; ++y)
                for (int z = 0; z < n; ++z) {
                    const uint32_t k = cell_key_hilbert(x, y, z, b);
                    if (k >= m || cx[k] != -1) { printf("bijection %%d\n", b); return 1; }
                    cx[k] = x; cy[k] = y; cz[k] = z;
                }
        for (size_t k = 0; k + 1 < m; ++k)
            if (abs(cx[k] - cx[k + 1]) + abs(cy[k] - cy[k + 1]) + abs(cz[k] - cz[k + 1]) != 1) {
                printf("adjacency %%d %%zu\n", b, k); return 1;
            }
        for (size_t g = 0; g < m; g += 8)
            for (int i = 1; i < 8; ++i)
                if ((cx[g + i] >> 1) != (cx[g] >> 1) || (cy[g + i] >> 1) != (cy[g] >> 1) ||
                    (cz[g + i] >> 1) != (cz[g] >> 1)) { printf("block %%d %%zu\n", b, g); return 1; }
    }
    printf("ok\n");
    return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_hilbert_key_is_a_face_adjacent_curve_of_aligned_blocks(tmp_path):
    """cell_key_hilbert is a bijection of the 2^b cube onto [0, 8^b), consecutive
    keys are face neighbours, and every aligned run of 8 keys is a 2x2x2 block
    (what the re-sort's dropped bits rely on)."""
    src = open(os.path.join(ROOT, "src_amd", "csrc", "common.hpp")).read()
    i = src.index("__host__ __device__ inline uint32_t cell_key_hilbert")
    j = src.index("\n}\n", i) + 3
    path = tmp_path / "hilbert.cpp"
    path.write_text(MAIN % src[i:j])
    exe = tmp_path / "hilbert"
    subprocess.run(["g++", "-O2", "-o", str(exe), str(path)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
