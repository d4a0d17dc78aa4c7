"""CPU check of the C++ drop-in for the sample/ interface
(level-set-method-code_amd/csrc/ls_sample_api.hpp): a port of the reference's
sample/main.cpp loop body written against the shim compiles for gfx950."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"

PORT = r"""
#include "ls_sample_api.hpp"
#include <cmath>
int main() {
    const int NX = 64, NY = 64;
    const double DX = 0.02 / (NX - 1), DY = 0.02 / (NY - 1), DT = 2.5e-6, SL = 0.4;
    Field2D G(NX, NY), u(NX, NY), v(NX, NY), G_initial(NX, NY), rhs(NX, NY);
    for (int j = 0; j < NY + 6; ++j)
        for (int i = 0; i < NX + 6; ++i) {
            G.data[j][i] = (j - 3) * DY - 0.005;
            G_initial.data[j][i] = G.data[j][i];
        }
    G.to_device(); G_initial.to_device(); u.fill(0.0); v.fill(0.8);
    G.apply_anchored_bc(G_initial);
    G.apply_periodic();
    compute_geqn_rhs(G, u, v, rhs, DX, DY, SL);
    RK3_step(G, u, v, DX, DY, DT, SL, G_initial);
    G.reinitialize(G_initial, DX, DY, 2);
    compute_first_order_upwind(G, u, v, DX, DY, DT, SL);
    save_G_to_bin(G, "/tmp/g.bin");
    return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc absent")
def test_sample_shim_port_compiles(tmp_path):
    src = tmp_path / "port.cpp"
    src.write_text(PORT)
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-std=c++17", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"),
                        "-I" + os.path.join(ROOT, "level-set-method-code_amd", "csrc"), str(src)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
