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
