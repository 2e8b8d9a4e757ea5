"""GPU tests of the files written from device fields (.vti, .bin) and the
refinement density remap, against the formats' layouts and the oracle."""
import os
import re
import struct
import sys

import numpy as np
import pytest

import oracle
from helpers import cantilever_arrays
from v119_2025_12_19_hexatop_aws_amd import hexa

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _solved(nx=6, ny=4, nz=3, h=(1.0, 0.5, 0.25)):
    fixed, f, rho = cantilever_arrays(nx, ny, nz, 9, lo=0.2)
    s = hexa.Solver(nx, ny, nz, h, E0=3.0, Emin=0.0, penal=1.0)
    s.set_problem(fixed, f)
    s.set_density(rho)
    u, reps, st = s.solve(tol=1e-10, precond=hexa.PRECOND_MG, semantics=hexa.SEMANTICS_CPU)
    assert st == hexa.OK
    return s, u[0], rho


def _parse_vti(path):
    raw = open(path, "rb").read()
    head, _, rest = raw.partition(b'<AppendedData encoding="raw">\n_')
    arrays = {}
    for name, off in re.findall(rb'Name="(\w+)"[^>]*offset="(\d+)"', head):
        o = int(off)
        nb = struct.unpack_from("<I", rest, o)[0]
        arrays[name.decode()] = np.frombuffer(rest, dtype="<f4", count=nb // 4, offset=o + 4)
    return head.decode(), arrays, rest


def test_write_vti_layout_and_fields(tmp_path):
    """ExportVTK.jl:16-111 layout: header, offsets, UInt32 byte counts, and the
    stress fields of the device solution."""
    say("vti")
    s, u, rho = _solved()
    E = 3.0
    st = s.stress(u, rho, E=E)
    path = str(tmp_path / "out")
    s.write_vti(path, E=E, origin=(0.0, 1.5, -2.0), write_vectors=True)
    head, arrs, rest = _parse_vti(path + ".vti")
    assert '<ImageData WholeExtent="0 6 0 4 0 3" Origin="0.0 1.5 -2.0" Spacing="1.0 0.5 0.25">' in head
    assert list(arrs) == ["Density", "L1_Stress", "VonMises", "PrincipalValues", "MaxPrincipalDirection",
                          "MinPrincipalDirection"]
    assert np.array_equal(arrs["Density"], rho.astype(np.float32))
    assert np.array_equal(arrs["L1_Stress"], st["l1"])
    assert np.array_equal(arrs["VonMises"], st["vonmises"])
    assert np.array_equal(arrs["PrincipalValues"], st["principal"].ravel())
    assert np.array_equal(arrs["MaxPrincipalDirection"], st["dir_max"].ravel())
    assert rest.endswith(b"\n  </AppendedData>\n</VTKFile>\n")
    s.write_vti(path + "_s.vti", E=E)
    head2, arrs2, _ = _parse_vti(path + "_s.vti")
    assert list(arrs2) == ["Density", "L1_Stress", "VonMises"] and 'offset="0"' in head2
    s.close()


def _read_bin(path):
    b = open(path, "rb").read()
    magic, ver, it, rad, thr, cnt, dx, dy, dz = struct.unpack_from("<IIiffIfff", b, 0)
    o = 36
    cen = np.frombuffer(b, "<f4", 3 * cnt, o); o += 12 * cnt
    den = np.frombuffer(b, "<f4", cnt, o); o += 4 * cnt
    sl1 = np.frombuffer(b, "<f4", cnt, o); o += 4 * cnt
    jl = struct.unpack_from("<I", b, o)[0]
    js = b[o + 4:o + 4 + jl].decode()
    assert o + 4 + jl == len(b)
    return dict(magic=magic, ver=ver, iter=it, radius=rad, threshold=thr, count=cnt, h=(dx, dy, dz),
                centroids=cen.reshape(-1, 3), density=den, signed_l1=sl1, json=js)


def test_write_bin_checkpoint(tmp_path):
    """Postprocessing.jl:370-471 layout; active elements (>= threshold),
    striding, signed l1; full_field writes every element."""
    say("bin")
    s, u, rho = _solved()
    E = 3.0
    st = s.stress(u, rho, E=E)
    d32 = rho.astype(np.float32)
    nx, ny, nz = 6, 4, 3
    hh = (1.0, 0.5, 0.25)
    s.write_bin(str(tmp_path / "a.bin"), E=E, iter=7, radius=2.5, threshold=0.5, json='{"k": 1}')
    r = _read_bin(str(tmp_path / "a.bin"))
    act = np.flatnonzero(d32 >= np.float32(0.5))
    assert (r["magic"], r["ver"], r["iter"], r["count"]) == (0x48455841, 2, 7, act.size)
    assert r["radius"] == np.float32(2.5) and r["h"] == hh and r["json"] == '{"k": 1}'
    assert np.array_equal(r["density"], d32[act])
    p = st["principal"][act]
    big = np.argmax(np.abs(p), axis=1)                  # first index of the largest |value|
    sgn = np.sign(p[np.arange(act.size), big])
    assert np.array_equal(r["signed_l1"], st["l1"][act] * sgn.astype(np.float32))
    i, j, k = act % nx, (act // nx) % ny, act // (nx * ny)
    assert np.allclose(r["centroids"], np.stack([(i + 0.5) * hh[0], (j + 0.5) * hh[1], (k + 0.5) * hh[2]], 1),
                       atol=1e-6)
    # strided export (max_export_cells) and the full field
    s.write_bin(str(tmp_path / "b.bin"), E=E, threshold=0.5, max_export_cells=5)
    rb = _read_bin(str(tmp_path / "b.bin"))
    step = act.size / 5
    want, cur = [], 1.0
    while cur <= act.size:
        want.append(act[int(np.floor(cur)) - 1]); cur += step
    assert np.array_equal(rb["density"], d32[want])
    s.write_bin(str(tmp_path / "c.bin"), E=E, threshold=0.5, full_field=True)
    rc = _read_bin(str(tmp_path / "c.bin"))
    assert rc["count"] == nx * ny * nz and np.array_equal(rc["density"], d32)
    s.close()


@pytest.mark.parametrize("hard", [False, True])
def test_regrid_matches_oracle_bitwise(hard):
    """MeshRefiner.jl:185-237 remap on the device equals the oracle's Float32
    restatement bit for bit (alpha: nearest-lower neighbour)."""
    say("regrid", hard)
    od, nd = (10, 6, 5), (17, 11, 9)
    L = (10.0, 6.0, 5.0)
    nh = tuple(L[a] / nd[a] for a in range(3))
    rng = np.random.default_rng(6)
    rho = rng.uniform(0.0, 1.0, int(np.prod(od))).astype(np.float32)
    alpha = rng.uniform(0.0, 1.0, rho.size).astype(np.float32)
    ro, ao = oracle.regrid(od, nd, nh, L, rho, hard=hard, alpha=alpha)
    rg, ag = hexa.regrid_density(od + (1.0, 1.0, 1.0), rho, nd + nh, L, hard=hard, alpha=alpha)
    assert np.array_equal(rg, ro)
    assert np.array_equal(ag, ao)
