"""Ensembles (nsp_ensemble_*) on the GPU: every member of a batched ensemble
step gives the same bits and corrector counts as the same pipe stepped alone
(same kernels and variants, launched per member), across two transform-plan
classes, both precisions and members that share operator data."""
import numpy as np
import pytest

import nsp

pytestmark = pytest.mark.gpu

A = dict(nr=8, nt=96, nz=513, lz=50.0, re=2000.0, dt=0.005)     # plans 144 / 768
B = dict(nr=12, nt=128, nz=1025, lz=50.0, re=2000.0, dt=0.005)  # plans 192 / 1536
SEED4 = 20261202 + 4 * 1000


def members():
    return [dict(A), dict(B), dict(A), dict(A, precision=nsp.MIXED), dict(B, re=2500.0)]


def seed_member(p, j):
    p.seeded_ic(SEED4 + j, 4, 4, 0, 0.02)


def test_ensemble_members_match_single_pipes():
    ms = members()
    print(f"[gpu] ensemble of {len(ms)}", flush=True)
    ens = nsp.Ensemble(ms)
    for j in range(len(ms)):
        seed_member(ens[j], j)
    n_e, err_e, fl_e = ens.step(3)
    assert np.all(fl_e == 0)
    for j, kw in enumerate(ms):
        p = nsp.Pipe(**kw)
        seed_member(p, j)
        n_p, err_p, fl_p = p.step(3)
        assert np.array_equal(n_p, n_e[:, j]), (j, n_p, n_e[:, j])
        assert np.array_equal(err_p, err_e[:, j]), j
        a, b = ens[j].get_state(), p.get_state()
        for c in range(3):
            assert np.array_equal(a[c], b[c]), (j, c)
        p.close()
    ens.close()


PAIR_VARIANT = dict(theta_inv=dict(alg=0, tl=8, nthr=256, pad=0, ord=1),
                    theta_fwd=dict(alg=1, tl=8, nthr=256, pad=0, ord=1),
                    z_alg=4, z_nthr=512, z_pad=1, z_pf=0, s_ord=1)


@pytest.mark.parametrize("z_alg", [3, 4])
def test_ensemble_pair_form_matches_single_pipes(z_alg):
    """The batched pair-form z pass (one or two rows per workgroup, rows of
    uneven members beyond a member's extent returning at entry): every member
    gives the bits and corrector counts of the same pipe stepped alone with the
    same kernel variants."""
    ms = members()
    var = dict(PAIR_VARIANT, z_alg=z_alg)
    print(f"[gpu] ensemble pair form z_alg {z_alg}", flush=True)
    ens = nsp.Ensemble(ms)
    for j in range(len(ms)):
        ens[j].set_tuning(var)  # the class's first member's choice is the class's
        seed_member(ens[j], j)
    n_e, err_e, fl_e = ens.step(3)
    assert np.all(fl_e == 0)
    for j, kw in enumerate(ms):
        p = nsp.Pipe(**kw)
        p.set_tuning(var)
        seed_member(p, j)
        n_p, err_p, _ = p.step(3)
        assert np.array_equal(n_p, n_e[:, j]), (j, n_p, n_e[:, j])
        assert np.array_equal(err_p, err_e[:, j]), j
        a, b = ens[j].get_state(), p.get_state()
        for c in range(3):
            assert np.array_equal(a[c], b[c]), (j, c)
        p.close()
    ens.close()


def test_ensemble_members_converge_independently():
    """A member with a tight tolerance would need more correctors; with a
    shared tol each member still stops on its own test (counts may differ)."""
    ms = [dict(A), dict(A)]
    ens = nsp.Ensemble(ms)
    seed_member(ens[0], 0)
    ens[1].seeded_ic(SEED4 + 1, 4, 4, 0, 0.3)  # strong perturbation: more correctors
    n_e, _, fl = ens.step(2)
    for j in range(2):
        p = nsp.Pipe(**ms[j])
        if j == 0:
            seed_member(p, 0)
        else:
            p.seeded_ic(SEED4 + 1, 4, 4, 0, 0.3)
        n_p, _, _ = p.step(2)
        assert np.array_equal(n_p, n_e[:, j])
        p.close()
    ens.close()


def test_ensemble_second_graph_matches_eager_pipes(monkeypatch):
    """With a tolerance that takes several correctors the ensemble step runs
    its second graph (correctors 2..maxit); every member still gives the bits,
    counts and flags of the same pipe stepped launch by launch."""
    ms = [dict(A, tol=1e-30, maxit=5), dict(A, tol=1e-30, maxit=5, precision=nsp.MIXED)]
    ens = nsp.Ensemble(ms)
    for j in range(len(ms)):
        seed_member(ens[j], j)
    n_e, err_e, fl_e = ens.step(2)
    assert n_e.max() > 1
    monkeypatch.setenv("NSP_NO_GRAPH", "1")
    for j, kw in enumerate(ms):
        p = nsp.Pipe(**kw)
        seed_member(p, j)
        n_p, err_p, fl_p = p.step(2)
        assert np.array_equal(n_p, n_e[:, j]), (j, n_p, n_e[:, j])
        assert np.array_equal(err_p, err_e[:, j]) and np.array_equal(fl_p, fl_e[:, j]), j
        a, b = ens[j].get_state(), p.get_state()
        for c in range(3):
            assert np.array_equal(a[c], b[c]), (j, c)
        p.close()
    ens.close()


def test_puff_ic_spectrum_is_a_gaussian_window():
    """Kind IC_PUFF: for m >= 1 the radial velocity of every axial mode l is the
    l = 0 coefficient times the spectrum of a Gaussian window of sigma 2.5R
    centred at z = LZ/2: exp(-(kz sigma)^2 / 2) exp(-i kz LZ/2)."""
    p = nsp.Pipe(**A)
    p.seeded_ic(SEED4, 4, 16, nsp.IC_PUFF, 0.07)
    ur, ut, uz = p.get_state()
    d = p.diagnostics()
    # analytically solenoidal: what is left is the radial stencil's truncation at NR = 8
    assert d["div_l2"] < 1e-3 * 0.07 and d["wall_max"] < 1e-14
    kz0 = 2 * np.pi / A["lz"]
    for m in range(1, 5):
        for l in range(1, 17):
            kz = kz0 * l
            env = np.exp(-0.5 * (kz * 2.5) ** 2) * np.exp(-1j * kz * A["lz"] / 2)
            np.testing.assert_allclose(ur[:, m, l], ur[:, m, 0] * env, rtol=1e-12, atol=1e-300)
            np.testing.assert_allclose(ur[:, A["nt"] - m, l], ur[:, A["nt"] - m, 0] * env,
                                       rtol=1e-12, atol=1e-300)
    for l in range(1, 17):  # m = 0: real amplitudes, ut carries the l = 0 shape
        kz = kz0 * l
        env = np.exp(-0.5 * (kz * 2.5) ** 2) * np.exp(-1j * kz * A["lz"] / 2)
        np.testing.assert_allclose(ut[:, 0, l], ut[:, 0, 0] * env, rtol=1e-12, atol=1e-300)
    assert np.abs(ur[:, :, 17:]).max() == 0.0
    # physical check: the axial energy profile of the m = 1 mode on the grid
    # z_j = j LZ / NZ peaks at LZ/2 and falls below 1e-3 of the peak by 4 sigma
    zj = np.arange(256) * A["lz"] / 256
    lv = np.arange(1, 17)
    i = A["nr"] // 2
    ph = np.exp(1j * np.outer(zj, kz0 * lv))
    # c(1, -l) is the conjugate of the stored c(-1, l)
    u_z = ur[i, 1, 0] + ph @ ur[i, 1, 1:17] + np.conj(ph) @ np.conj(ur[i, A["nt"] - 1, 1:17])
    e_z = np.abs(u_z) ** 2
    assert abs(zj[np.argmax(e_z)] - A["lz"] / 2) < 0.5
    far = np.abs(zj - A["lz"] / 2) > 4 * 2.5
    assert e_z[far].max() < 1e-3 * e_z.max()
    p.close()


def test_additive_ic_adds_only_the_perturbation():
    """IC_ADD adds the seeded perturbation to the current state (no base flow):
    (puff + add(seed s)) - puff == flat(seed s) - flat(rms 0)."""
    pipes = [nsp.Pipe(**A) for _ in range(4)]
    pipes[0].seeded_ic(SEED4, 4, 16, nsp.IC_PUFF, 0.07)
    pipes[0].seeded_ic(SEED4 + 7, 4, 4, nsp.IC_FLAT | nsp.IC_ADD, 1e-3)
    pipes[1].seeded_ic(SEED4, 4, 16, nsp.IC_PUFF, 0.07)
    pipes[2].seeded_ic(SEED4 + 7, 4, 4, nsp.IC_FLAT, 1e-3)
    pipes[3].seeded_ic(SEED4 + 7, 4, 4, nsp.IC_FLAT, 0.0)
    s = [q.get_state() for q in pipes]
    for c in range(3):
        d1 = s[0][c] - s[1][c]
        d2 = s[2][c] - s[3][c]
        assert np.abs(d2).max() > 0
        np.testing.assert_allclose(d1, d2, rtol=0, atol=1e-15)
    for q in pipes:
        q.close()


def test_ensemble_members_match_oracle():
    """The batched ensemble step (mod_integrate.cu:112-131 per member) against
    the CPU oracle: uneven members at the config-4 transform plans (144/768 and
    192/1536), two radial sizes, two Reynolds numbers, both precisions, each
    started from the config-4 recipe (localized puff plus a member-specific
    perturbation).  Three steps; FP64 members within 1e-12 relative L2 of the
    oracle on every field, mixed members within 1e-6, identical corrector
    counts."""
    from oracle_lib import Oracle
    ms = [dict(A), dict(B, nr=8), dict(A, nr=12, re=1800.0), dict(A, precision=nsp.MIXED),
          dict(B, precision=nsp.MIXED)]
    print(f"[gpu] ensemble of {len(ms)} against the oracle", flush=True)
    ens = nsp.Ensemble(ms)
    for j in range(len(ms)):
        ens[j].seeded_ic(SEED4, 4, 16, nsp.IC_PUFF, 0.1)
        ens[j].seeded_ic(SEED4 + 1 + j, 4, 4, nsp.IC_FLAT | nsp.IC_ADD, 1e-3)
    u0 = [ens[j].get_state() for j in range(len(ms))]
    n_e, _, fl_e = ens.step(3)
    assert np.all(fl_e == 0)
    for j, kw in enumerate(ms):
        shape = {k: v for k, v in kw.items() if k != "precision"}
        o = Oracle(**shape)
        o.set_state(*u0[j])
        n_o, _ = o.step(3)
        assert np.array_equal(n_o, n_e[:, j]), (j, n_o, n_e[:, j])
        tol = 1e-6 if kw.get("precision") == nsp.MIXED else 1e-12
        a, b = ens[j].get_state(), o.get_state()
        errs = [np.linalg.norm(x - y) / np.linalg.norm(y) for x, y in zip(a, b)]
        print(f"[gpu] member {j} {kw}: n_it {n_o.tolist()} rel L2 {errs}", flush=True)
        assert max(errs) <= tol, (j, errs)
        del o
    ens.close()


def test_config4_full_size_ensemble_properties():
    """Config 4 at full size: 32 pipes at (48, 96, 513) and 32 at
    (64, 128, 1025), Re 2000, L 50R, dt 0.005, mixed form, each from the
    localized puff plus a member-specific 1e-6 perturbation.  Three steps of
    the batched ensemble: no flags, bulk flux 0.5 to 1e-12, no-slip to 1e-12,
    and every member's kinetic energy within 1% of its start."""
    cfg_a = dict(nr=48, nt=96, nz=513, lz=50.0, re=2000.0, dt=0.005, precision=nsp.MIXED)
    cfg_b = dict(nr=64, nt=128, nz=1025, lz=50.0, re=2000.0, dt=0.005, precision=nsp.MIXED)
    ms = [cfg_a] * 32 + [cfg_b] * 32
    print("[gpu] config-4 ensemble, 64 members", flush=True)
    ens = nsp.Ensemble(ms)
    puff_rms = 0.5 * 0.5 * np.sqrt(np.sqrt(np.pi) * 2.5 / 50.0)
    for j in range(len(ms)):
        ens[j].seeded_ic(SEED4, 4, 16, nsp.IC_PUFF, puff_rms)
        ens[j].seeded_ic(SEED4 + 1 + j, 4, 4, nsp.IC_FLAT | nsp.IC_ADD, 1e-6 * 0.5)
    ke0 = [ens[j].diagnostics()["ke"] for j in range(len(ms))]
    n_it, _, flags = ens.step(3)
    print(f"[gpu] config-4 ensemble n_it per step {n_it.max(axis=1).tolist()}", flush=True)
    assert np.all(flags == 0)
    for j in range(len(ms)):
        d = ens[j].diagnostics()
        assert abs(d["flux"] - 0.5) < 1e-12, j
        assert d["wall_max"] < 1e-12, j
        assert abs(d["ke"] / ke0[j] - 1) < 1e-2, j
    ens.close()
