"""State conversion and output on the GPU against the oracle: remesh
(nsp_remesh vs the oracle's restatement of tools/Re_Mesh/src/mod_remesh.cu)."""
import numpy as np
import pytest

import nsp
import oracle_lib

pytestmark = pytest.mark.gpu

BASE = dict(lz=2 * np.pi, re=1000.0, dt=0.01)


def random_state(p, seed):
    rng = np.random.default_rng(seed)
    f = [rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape) for _ in range(3)]
    p.set_state(*f)
    return f


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# (old NR, NT, NZ) -> (new NR, NT, NZ): refine, coarsen, mixed, unchanged
REMESH_CASES = [((16, 8, 9), (24, 12, 17)), ((24, 12, 17), (16, 8, 9)), ((12, 16, 9), (20, 8, 13)),
                ((16, 8, 9), (16, 8, 9)), ((48, 96, 33), (64, 128, 49))]


@pytest.mark.parametrize("old,new", REMESH_CASES)
@pytest.mark.parametrize("flags", [nsp.REMESH_REFERENCE, nsp.REMESH_MIRROR])
def test_remesh_matches_oracle(old, new, flags):
    print(f"[gpu] remesh {old} -> {new} flags {flags}", flush=True)
    src = nsp.Pipe(*old, **BASE)
    dst = nsp.Pipe(*new, **BASE)
    f = random_state(src, 100 + sum(old))
    random_state(dst, 7)  # every entry of the new state must be overwritten
    nsp.remesh(dst, src, flags)
    got = dst.get_state()
    want = oracle_lib.remesh(f, new, flags)
    for c in range(3):
        assert rel_l2(got[c], want[c]) <= 1e-12, c
        # modes absent from the old grid are exactly zero
        assert np.all((want[c] == 0) <= (got[c] == 0)), c
    src.close()
    dst.close()


def test_remesh_state_then_steps():
    """A seeded state remeshed to a finer grid steps: finite, correctors
    converge after the first step (the interpolated field is not projected),
    bulk flux held at Qdes."""
    src = nsp.Pipe(16, 16, 17, **BASE)
    src.seeded_ic(5, 3, 3, nsp.IC_FLAT, 0.02)
    dst = nsp.Pipe(24, 24, 25, **BASE)
    nsp.remesh(dst, src, nsp.REMESH_MIRROR)
    n_it, _, flags = dst.step(5)
    assert np.all((flags & nsp.FLAG_NAN) == 0) and np.all(flags[1:] == 0) and n_it[-1] <= 4
    d = dst.diagnostics()
    assert np.isfinite(d["ke"]) and abs(d["flux"] - 0.5) < 1e-12
    src.close()
    dst.close()


def test_remesh_rejects_same_handle():
    p = nsp.Pipe(16, 8, 9, **BASE)
    with pytest.raises(nsp.NspError):
        nsp.remesh(p, p)
    p.close()


# ---------------------------------------------------------------- physical space
OUT_SHAPES = [dict(nr=16, nt=8, nz=9, lz=2 * np.pi, re=500.0, dt=0.01),
              dict(nr=8, nt=96, nz=513, lz=50.0, re=2000.0, dt=0.005),
              dict(nr=10, nt=64, nz=65, lz=4 * np.pi, re=2000.0, dt=0.005)]


def paired(shape, precision, seed):
    p = nsp.Pipe(**shape, precision=precision)
    o = oracle_lib.Oracle(**shape)
    f = random_state(p, seed)
    o.set_state(*f)
    return p, o


# the mixed form exists only for sizes with compiled transform plans
@pytest.mark.parametrize("shape,precision", [(s, nsp.FP64) for s in OUT_SHAPES] +
                         [(s, nsp.MIXED) for s in OUT_SHAPES[1:]])
def test_physical_field_matches_oracle(shape, precision):
    print(f"[gpu] physical field {shape} precision {precision}", flush=True)
    p, o = paired(shape, precision, 21)
    for c in range(3):
        got, want = p.physical(c), o.physical(c)
        assert got.shape == want.shape
        assert rel_l2(got, want) <= 1e-12, c
        assert np.all(got[:, :, -2:] == 0)
    p.close()


@pytest.mark.parametrize("shape", OUT_SHAPES)
def test_cross_sections_and_mean_profile_match_oracle(shape):
    print(f"[gpu] cross sections {shape}", flush=True)
    p, o = paired(shape, nsp.FP64, 22)
    q, u2 = p.cross_sections()
    qo, u2o = o.cross()
    assert rel_l2(q, qo) <= 1e-12 and rel_l2(u2, u2o) <= 1e-12
    assert np.array_equal(p.mean_profile(), p.get_state()[2][:, 0, 0].real)
    p.close()


def read_rows(path):
    with open(path) as f:
        return [line.rstrip("\n") for line in f]


def test_io_files_in_reference_format(tmp_path):
    shape = OUT_SHAPES[0]
    p, o = paired(shape, nsp.FP64, 23)
    p.io_open(tmp_path)
    p.io_write(0.25)
    d = o.diagnostics()
    q, u2 = o.cross()
    p.io_close()
    num = r"-?\d\.\d{8}e[+-]\d{2}"
    import re
    fr = read_rows(tmp_path / "io_friction.txt")
    assert len(fr) == 1 and re.fullmatch(f"{num}  {num}  {num}", fr[0])
    np.testing.assert_allclose([float(x) for x in fr[0].split()], [0.25, d["u_cl"], d["u_tau"]], rtol=1e-7)
    nzf = 3 * (shape["nz"] - 1)
    for name, want in (("io_qcross.txt", q), ("io_ucross.txt", u2)):
        rows = read_rows(tmp_path / name)
        assert len(rows) == 2
        assert re.fullmatch(f"(?:{num}  ){{{nzf + 1}}}", rows[0]) and re.fullmatch(f"(?:{num}  ){{{nzf + 1}}}", rows[1])
        head = np.array([float(x) for x in rows[0].split()])
        np.testing.assert_allclose(head, np.r_[0.0, shape["lz"] * np.arange(nzf) / nzf], rtol=1e-8, atol=1e-300)
        vals = np.array([float(x) for x in rows[1].split()])
        np.testing.assert_allclose(vals, np.r_[0.25, want], rtol=1e-7)
    mp = read_rows(tmp_path / "io_meanpr.txt")
    r = p.tables()["r"]
    np.testing.assert_allclose([float(x) for x in mp[0].split()], np.r_[0.0, r], rtol=1e-8, atol=1e-300)
    np.testing.assert_allclose([float(x) for x in mp[1].split()], np.r_[0.25, p.get_state()[2][:, 0, 0].real],
                               rtol=1e-7, atol=1e-300)
    p.close()


# ---------------------------------------------------------------- restart metadata
def test_restart_directory_round_trip(tmp_path):
    shape = OUT_SHAPES[0]
    p = nsp.Pipe(**shape)
    p.seeded_ic(9, 3, 3, nsp.IC_FLAT, 0.05)
    p.step(3)
    p.write_restart(tmp_path, t=0.03, step=3)
    saved = p.get_state()
    meta = dict(line.split() for line in open(tmp_path / "restart.txt"))
    assert meta["format"] == "nspipe-restart-1" and int(meta["nr"]) == shape["nr"] and float(meta["time"]) == 0.03
    q = nsp.Pipe(**shape)
    assert q.read_restart(tmp_path) == (0.03, 3)
    for a, b in zip(p.get_state(), q.get_state()):
        assert np.array_equal(a, b)
    # continuing from the restart gives the same steps as continuing the original
    p.step(2)
    q.step(2)
    for a, b in zip(p.get_state(), q.get_state()):
        assert np.array_equal(a, b)
    # without the metadata the directory is refused (the raw files carry no
    # grid); they still read as plain reference state files
    (tmp_path / "restart.txt").unlink()
    with pytest.raises(nsp.NspError):
        q.read_restart(tmp_path)
    q.read_state_files(*(str(tmp_path / f"u{c}.bin") for c in "rtz"))
    for a, b in zip(saved, q.get_state()):
        assert np.array_equal(a, b)
    # another grid is refused
    p.write_restart(tmp_path, t=1.0, step=100)
    r = nsp.Pipe(shape["nr"] + 2, shape["nt"], shape["nz"], shape["lz"], shape["re"], shape["dt"])
    with pytest.raises(nsp.NspError):
        r.read_restart(tmp_path)
    for x in (p, q, r):
        x.close()


def test_driver_writes_diagnostics_and_restart(tmp_path):
    """nsp_run (the replacement of main.cu): diagnostics rows at the --dt_*
    cadence, a restart directory at the end that a second run continues."""
    import subprocess
    import os
    exe = os.path.join(os.path.dirname(nsp.LIB_PATH), "nsp_run")
    io1, save1 = tmp_path / "io", tmp_path / "save"
    io1.mkdir()
    save1.mkdir()
    args = [exe, "--nr", "16", "--nt", "8", "--nz", "9", "--lz", "6.283185307179586", "--re", "500", "--dt", "0.01",
            "--seed", "4", "--rms", "0.05", "--out_every", "0"]
    print("[gpu] nsp_run first leg", flush=True)
    subprocess.run(args + ["--steps", "20", "--io_dir", str(io1), "--save_dir", str(save1), "--dt_frc", "0.049",
                           "--dt_qcr", "0.099", "--dt_mnp", "0.049"], check=True, timeout=120, capture_output=True)
    fr = read_rows(io1 / "io_friction.txt")
    assert len(fr) == 4 and abs(float(fr[-1].split()[0]) - 0.2) < 1e-12
    assert len(read_rows(io1 / "io_qcross.txt")) == 1 + 2 and len(read_rows(io1 / "io_meanpr.txt")) == 1 + 4
    meta = dict(line.split() for line in open(save1 / "restart.txt"))
    assert int(meta["step"]) == 20
    print("[gpu] nsp_run second leg", flush=True)
    out = subprocess.run([exe, "--nr", "16", "--nt", "8", "--nz", "9", "--lz", "6.283185307179586", "--re", "500",
                          "--dt", "0.01", "--steps", "5", "--restart_dir", str(save1), "--save_dir", str(save1),
                          "--out_every", "0"], check=True, timeout=120, capture_output=True, text=True)
    assert "restart at t=2.00000000e-01, step 20" in out.stdout
    meta = dict(line.split() for line in open(save1 / "restart.txt"))
    assert int(meta["step"]) == 25 and abs(float(meta["time"]) - 0.25) < 1e-12
