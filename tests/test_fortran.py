"""The Fortran boundary: the iso_c_binding interface module generated from
include/gemini_fk.h is committed in sync with the header, compiles with
amdflang, links against libgemini_fk.so, and a Fortran program computes one
frequency's Green-FK columns through it (GPU test, compared with the oracle).
The drop-in toroidalMinors / spheroidalMinors modules serve the reference's
per-degree calls from GPU batches (GPU test against the oracle's minors)."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "gemini_unified_amd", "fortran")
PROG = os.path.join(ROOT, "tests", "fortran", "gfk_call")
SHIM = os.path.join(ROOT, "tests", "fortran", "minors_shim")
FLANG = os.environ.get("AMDFLANG", "/opt/rocm/bin/amdflang")
needs_flang = pytest.mark.skipif(not os.path.exists(FLANG) and not os.path.exists(PROG),
                                 reason="amdflang not installed")


def _program(path=PROG):
    """A test program, built only when it is not there yet."""
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "fortran")], check=True)
    return path


def test_interface_module_matches_header():
    r = subprocess.run([sys.executable, os.path.join(FDIR, "gen_interfaces.py"), "--check"])
    assert r.returncode == 0, "geminiFKc.f90 is stale: run gemini_unified_amd/fortran/gen_interfaces.py"


def test_interface_covers_every_entry_point():
    sys.path.insert(0, FDIR)
    import gen_interfaces
    from gemini_unified_amd import _lib
    with open(gen_interfaces.HEADER) as fh:
        _, protos = gen_interfaces.parse(fh.read())
    assert sorted(n for n, _ in protos) == sorted(_lib.EXPORTS)


@needs_flang
def test_fortran_program_links_and_creates_a_context():
    out = subprocess.run([_program(), "create"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    line = [ln for ln in out.stdout.splitlines() if ln.startswith("gm_ctx_create status")][0]
    rc = int(line.split()[-1])
    # 0 on a GPU host; without a device the library's own error code and message
    assert rc == 0 or (rc == 10 and "no HIP device" in out.stdout), out.stdout


@pytest.mark.gpu
@needs_flang
def test_fortran_gfk_frequency_matches_oracle(tmp_path, c1_setup):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from gemini_unified_amd import MC_PI
    nem, nodes, jre = c1_setup
    f, nwn, sigma, eps = 44 / 2048.0, 40, 5.0 / 2048.0, 1e-6
    lines = [f"{nem.nk} {nem.nlay} {nem.nic} {nem.noc} {nem.fsflag} {nem.equidis}",
             f"{nem.rearth!r} {nem.ric!r} {nem.roc!r} {nem.fref!r}"]
    for i in range(nem.nk):
        lines.append(" ".join(repr(float(a[i])) for a in (nem.rk, nem.ro, nem.vpv, nem.vph, nem.vsv, nem.vsh,
                                                          nem.eta, nem.qkinv, nem.qminv)))
    for i in range(nem.nlay):
        lines.append(f"{int(nem.iktop[i])} {float(nem.rb[i])!r} {int(nem.iflso[i])}")
    lines.append(str(nodes.nnod))
    for i in range(nodes.nnod):
        lines.append(f"{float(nodes.rnod[i])!r} {int(nodes.state[i])} {int(nodes.layer[i])}")
    lines.append(f"{f!r} {nwn} {sigma!r} {eps!r} 3 {jre} 1")
    fin, fout = tmp_path / "in.txt", tmp_path / "out.bin"
    fin.write_text("\n".join(lines) + "\n")
    print(f"Fortran gm_gfk_batch: f = {f}, l = 0..{nwn - 1}", flush=True)
    r = subprocess.run([_program(), "gfk", str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    rc, n, nsp, nn = np.frombuffer(raw[:16], np.int32)
    assert rc == 0 and (n, nsp, nn) == (nwn, 10, nodes.nnod)
    m = n * nsp * nn
    gfre = np.frombuffer(raw[16:16 + 4 * m], np.float32).reshape(nn, nsp, n)
    gfim = np.frombuffer(raw[16 + 4 * m:16 + 8 * m], np.float32).reshape(nn, nsp, n)
    status = np.frombuffer(raw[16 + 8 * m:16 + 8 * m + 4 * n], np.int32)
    assert np.all(status == 0)
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)
    rc, ref, st = orc.gfk_degrees(2 * MC_PI * f, -sigma, 0, nwn, jre, 1, eps)
    assert rc == 0
    # the exact variant: the file's FP32 values are the oracle's FP64 columns rounded
    np.testing.assert_array_equal(gfre, ref.real.astype(np.float32))
    np.testing.assert_array_equal(gfim, ref.imag.astype(np.float32))


@needs_flang
def test_minors_shims_refuse_calls_before_init():
    """computeToroidalMinors before gpuMinorsInit: a level-2 error in errmsg, as the reference reports
    its own failures (toroidalMinors.f90:90-97)."""
    out = subprocess.run([_program(SHIM), "nocontext"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "level 2: gpuMinorsInit was not called" in out.stdout


def _model_input(nem, nodes, f, nwn, sigma, eps, jre):
    lines = [f"{nem.nk} {nem.nlay} {nem.nic} {nem.noc} {nem.fsflag} {nem.equidis}",
             f"{nem.rearth!r} {nem.ric!r} {nem.roc!r} {nem.fref!r}"]
    for i in range(nem.nk):
        lines.append(" ".join(repr(float(a[i])) for a in (nem.rk, nem.ro, nem.vpv, nem.vph, nem.vsv, nem.vsh,
                                                          nem.eta, nem.qkinv, nem.qminv)))
    for i in range(nem.nlay):
        lines.append(f"{int(nem.iktop[i])} {float(nem.rb[i])!r} {int(nem.iflso[i])}")
    lines.append(str(nodes.nnod))
    for i in range(nodes.nnod):
        lines.append(f"{float(nodes.rnod[i])!r} {int(nodes.state[i])} {int(nodes.layer[i])}")
    lines.append(f"{f!r} {nwn} {sigma!r} {eps!r} 3 {jre} 1")
    return "\n".join(lines) + "\n"


@pytest.mark.gpu
@needs_flang
def test_minors_shims_serve_the_degree_loop(tmp_path):
    """The reference's degree loop calls computeToroidalMinors, computeSpheroidalMinors and
    computeUnitJumpGreenFunctions per degree; the drop-in modules fill toromin%u_tor, d_tor, det,
    spheromin%u_minors, d_minors, det, the minors' matrices and ujgf%ujgfsph, ujgftor from GPU
    batches of 16 degrees (exact variant), equal to the oracle's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from gemini_unified_amd import MC_PI, ExternalRadialNodes, dll1_of, global_model
    nem = global_model(1, 200)
    prn = np.array([600e3, 2000e3, 5000e3, nem.rearth - 300e3, nem.rearth - 30e3, nem.rearth])
    nodes = ExternalRadialNodes.split_at_discos(nem, prn, nem.rearth)
    jre = nodes.node_index(nem.rearth - 30e3)
    f, nwn, sigma, eps = 30 / 2048.0, 40, 5.0 / 2048.0, 1e-6
    fin, fout = tmp_path / "in.txt", tmp_path / "out.bin"
    fin.write_text(_model_input(nem, nodes, f, nwn, sigma, eps, jre))
    print(f"Fortran minors shims: f = {f}, l = 0..{nwn - 1}", flush=True)
    r = subprocess.run([_program(SHIM), "minors", str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw, pos = fout.read_bytes(), 0
    nn = nodes.nnod
    orc = oracle.Oracle(nem, nodes)
    orc.set_frequency(f, 3)

    def take(count, dtype):
        nonlocal pos
        a = np.frombuffer(raw[pos:pos + count * np.dtype(dtype).itemsize], dtype)
        pos += count * np.dtype(dtype).itemsize
        return a

    idx = {(0, 1): 0, (0, 2): 1, (0, 3): 2, (1, 2): 3, (1, 3): 4, (2, 3): 5}
    for l in range(nwn):
        assert take(1, np.int32)[0] == 0
        ut = take(2 * nn, np.complex128).reshape(nn, 2)
        dt = take(2 * nn, np.complex128).reshape(nn, 2)
        dett = take(1, np.complex128)[0]
        us = take(6 * nn, np.complex128).reshape(nn, 6)
        ds = take(6 * nn, np.complex128).reshape(nn, 6)
        dets = take(1, np.complex128)[0]
        mtil = take(16 * nn, np.complex128).reshape(nn, 4, 4).transpose(0, 2, 1)     # Fortran (4, 4, nnod)
        meps = take(16 * nn, np.complex128).reshape(nn, 4, 4).transpose(0, 2, 1)
        ujs = take(16 * nn, np.complex128).reshape(nn, 4, 4)                         # Fortran (4, njs, nnod)
        ujt = take(4 * nn, np.complex128).reshape(nn, 2, 2)
        om, dll1 = 2 * MC_PI * f, dll1_of(l, nem.rearth)
        rc, u, d, det = orc.toroidal_minors(om, -sigma, dll1, jre, jre, eps)
        assert rc == 0
        np.testing.assert_allclose(ut, u, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(dt, d, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(dett, det[jre - 1], rtol=1e-12, atol=1e-300)
        rc, u, d, ulg, dlg, det = orc.spheroidal_minors(om, -sigma, dll1, jre, jre, eps)
        assert rc == 0
        np.testing.assert_allclose(us, u, rtol=1e-12, atol=1e-12 * np.abs(u).max())
        np.testing.assert_allclose(ds, d, rtol=1e-12, atol=1e-12 * np.abs(d).max())
        np.testing.assert_allclose(dets, det[jre - 1], rtol=1e-11)
        # the minors' matrices (spheroidalMinors.f90:306-316): solid nodes, l > 0 (zero in the fluid core)
        for n in range(nn):
            want_t, want_e = np.zeros((4, 4), complex), np.zeros((4, 4), complex)
            if l > 0 and nodes.state[n] == 0:
                for (i, j), m in idx.items():
                    vf = -1.0 + 2.0 * ((i + j + 2) % 2)
                    want_t[i, j], want_t[j, i] = u[n, m], -u[n, m]
                    want_e[i, j], want_e[j, i] = vf * d[n, 5 - m], -vf * d[n, 5 - m]
            np.testing.assert_allclose(mtil[n], want_t, rtol=1e-12, atol=1e-12 * np.abs(u).max())
            np.testing.assert_allclose(meps[n], want_e, rtol=1e-12, atol=1e-12 * np.abs(d).max())
        # computeUnitJumpGreenFunctions: ujgfsph(comp, jump, node), ujgftor, as the oracle's column
        rc, gf, ref_s, ref_t, flag = orc.gfk_column(om, -sigma, dll1, jre, 1, eps)
        assert rc == 0
        np.testing.assert_allclose(ujs, ref_s, rtol=1e-11, atol=1e-11 * np.abs(ref_s).max())
        np.testing.assert_allclose(ujt, ref_t, rtol=1e-11, atol=1e-11 * max(np.abs(ref_t).max(), 1e-300))
    assert pos == len(raw)
