"""The C-ABI library loads, exports every declared entry point, and its
host-only set-up helpers agree with the oracle (no GPU needed)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from cetasim_amd import capi, configs

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("cetasim_hip.h", "cetasim_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(cts_[a-z0-9_]+)\s*\(", text))
    return names


def test_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_functions()
    assert len(names) >= 30
    for n in sorted(names):
        assert hasattr(lib, n), n
    # the Python bindings cover the same set
    assert names == set(capi.SIGNATURES), names ^ set(capi.SIGNATURES)


def test_struct_layout_matches_headers(tmp_path):
    """sizeof/offsetof of every struct as the C compiler lays it out."""
    structs = {"cts_ring": capi.Ring, "cts_section": capi.Section, "cts_moments": capi.Moments,
               "cts_derived": capi.Derived, "cts_twiss": capi.Twiss, "cts_ring_input": capi.RingInput,
               "cts_ring_derived": capi.RingDerived, "cts_ion_config": capi.IonConfig}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cetasim_host.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)])
    c = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
    for cname, py in structs.items():
        assert int(c[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(c[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def twiss_pair():
    t1 = dict(beta_x=8.0, alpha_x=0.7, eta_x=0.05, etap_x=-0.01, beta_y=12.0, alpha_y=-0.4,
              eta_y=0.003, etap_y=0.002, beta_z=0.57)
    t2 = dict(beta_x=11.0, alpha_x=-0.3, eta_x=0.0, etap_x=0.004, beta_y=6.0, alpha_y=0.2,
              eta_y=-0.001, etap_y=0.0, beta_z=0.57)
    return t1, t2


def test_build_section_matches_oracle_bitwise():
    lib = capi.load()
    t1, t2 = twiss_pair()
    sec = capi.Section()
    assert lib.cts_build_section(capi.Twiss(**t1), capi.Twiss(**t2), 1.1, 1.7, sec) == 0
    b1h1, h1b1, h2b2 = O.section_matrices(O.OrcTwiss(**t1), O.OrcTwiss(**t2))
    assert np.array_equal(np.array(sec.b1h1[:]), b1h1)
    assert np.array_equal(np.array(sec.h1b1[:]), h1b1)
    assert np.array_equal(np.array(sec.h2b2[:]), h2b2)


def test_section_map_closed_forms():
    """B1H1 * InvH1InvB1 = 1; with eta = 0 the one-section map is a rotation
    by 2*pi*nu in normalized coordinates (det 1, eigenvalues exp(+-2 pi i nu))
    and symplectic."""
    t1, _ = twiss_pair()
    b1h1, h1b1, h2b2 = [m.reshape(6, 6) for m in O.section_matrices(O.OrcTwiss(**t1), O.OrcTwiss(**t1))]
    assert np.allclose(b1h1 @ h1b1, np.eye(6), atol=1e-14)
    assert np.array_equal(h1b1, h2b2)  # t2 = t1
    nux, nuy = 0.18, 0.27
    R = np.eye(6)
    for k, nu in ((0, nux), (2, nuy)):
        c, s = math.cos(2 * math.pi * nu), math.sin(2 * math.pi * nu)
        R[k:k + 2, k:k + 2] = [[c, s], [-s, c]]
    M = h2b2 @ R @ b1h1
    J = np.zeros((6, 6))
    for k in (0, 2, 4):
        J[k, k + 1], J[k + 1, k] = 1, -1
    assert np.allclose(M.T @ J @ M, J, atol=1e-12)
    ev = np.linalg.eigvals(M[:2, :2] if t1["eta_x"] == 0 else M)
    tunes = sorted(set(np.round(np.abs(np.angle(ev)) / (2 * math.pi), 12)))
    assert any(abs(t - nux) < 1e-10 for t in tunes) and any(abs(t - nuy) < 1e-10 for t in tunes)


def test_derive_ring():
    lib = capi.load()
    ring = capi.Ring()
    der = capi.RingDerived()
    ri = configs.ring_input()
    assert lib.cts_derive_ring(ri, ring, der) == 0
    # T0 = C/(c beta) = 2 us, Qs = 0.005 by construction of the RF voltage
    assert ring.t0 == pytest.approx(2e-6, rel=1e-12)
    assert der.workqz == pytest.approx(0.005, rel=1e-12)
    # single cavity: Re(Vc) = U0 (src/CavityResonator.cpp:121-125)
    assert ring.rf_vr[0] == pytest.approx(ri.u0, rel=1e-12)
    assert ring.emit_nat[0] == pytest.approx(4e-11, rel=1e-12)
    assert ring.emit_nat[1] == pytest.approx(1.6e-12, rel=1e-12)
