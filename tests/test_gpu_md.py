"""GPU parity of the MD hot path (libac4.so through the C ABI) against the
oracle: bit-exact cells and neighbour counts, FP64 forces within 1e-12,
trajectories, shuffle rule, batch independence, graph replay."""
import contextlib

import numpy as np
import pytest

import oracle_lib as OL
from test_oracle_md import lattice_state, _rel_force_err

pytestmark = pytest.mark.gpu


def make_md(ac4, states, f0=1.5, T=0.6, tau=0, prec=0, seeds=None):
    ps = []
    for k, (s, L) in enumerate(states):
        ps.append(ac4.ReplicaParams(len(s["x"]), 1.2, T, f0, -1, tau, (seeds or {}).get(k, 1000 + k)))
    md = ac4.MD(ps, precision=prec)
    for k, (s, L) in enumerate(states):
        md.set_state(k, s["x"], s["y"], s["z"], s["xu"], s["yu"], s["zu"], s["px"], s["py"], s["pz"], s["type"],
                     s["act_tag"], s["kx"], s["ky"], s["kz"])
    return md


@contextlib.contextmanager
def tuning(ac4, tpp, nlist_mode=3):
    """Kernel shape for handles created inside the block (default restored)."""
    ac4.set_tuning(tpp, nlist_mode)
    try:
        yield
    finally:
        ac4.set_tuning(0, 3)
        ac4.set_skin(0.35)


_RELAXED = {}


def relaxed_state(ac4, n, seed, T=0.6, f0=1.5):
    """Lattice start brought to temperature on the GPU (ac4_md_relax), as the
    reference does with coldstart/breakxal before Nose-Hoover dynamics."""
    key = (n, seed, T, f0)
    if key not in _RELAXED:
        # unjittered sites (spacing ~0.9 sigma): a jittered lattice can put a
        # pair close enough for the first kick to blow the run up
        s, L = lattice_state(n, T=T, seed=seed, jitter=0.0)
        md = make_md(ac4, [(s, L)], f0=f0, T=T)
        md.relax(300, 1.5, 1.0)
        g = md.get_state(0)
        st = dict(s)
        st.update({k: g[k] for k in ("x", "y", "z", "xu", "yu", "zu", "px", "py", "pz")})
        md.close()
        _RELAXED[key] = (st, L)
    return _RELAXED[key]


def make_relaxed_md(ac4, specs, tau, prec, f0=1.5, T=0.6):
    """specs: list of (n, state seed, rng seed)."""
    states = [relaxed_state(ac4, n, sd, T, f0) for n, sd, _ in specs]
    md = make_md(ac4, states, f0=f0, T=T, tau=tau, prec=prec, seeds={k: rs for k, (_, _, rs) in enumerate(specs)})
    for k, (n, sd, rs) in enumerate(specs):
        md.init_chain(k, seed=sd)
    return md


def test_relax_reaches_target_temperature(gpu):
    s, L = lattice_state(1500, T=0.5, seed=8)
    md = make_md(gpu, [(s, L)], f0=1.0, T=0.5)
    md.relax(200, 2.0, 1.0)
    st = md.stats(0)
    assert st["T"] == pytest.approx(0.5, rel=1e-12) and st["steps"] == 0
    md.init_chain(0)
    md.step(200)
    md.sync()
    assert 0.3 < md.stats(0)["T"] < 0.8


@pytest.mark.parametrize("n,jit", [(1000, 0.05), (1000, 0.4), (2000, 0.05), (10000, 0.1), (875, 0.2)])
def test_cells_and_full_counts_bit_exact(gpu, n, jit):
    s, L = lattice_state(n, jitter=jit, seed=n + 1)
    md = make_md(gpu, [(s, L)])
    cell, cnt = md.update_cell_list(0)
    _, _, ocell = OL.half_list(s["x"], s["y"], s["z"], L)
    assert np.array_equal(cell, ocell)
    assert np.array_equal(cnt, OL.full_counts(s["x"], s["y"], s["z"], L))


@pytest.mark.parametrize("n,f0", [(1000, 1.5), (2000, 0.0), (4000, 2.0)])
def test_forces_fp64_match_oracle(gpu, n, f0):
    s, L = lattice_state(n, jitter=0.06, seed=n + 7)
    md = make_md(gpu, [(s, L)], f0=f0)
    pv = md.calculate_forces()[0]
    cnt, lst, _ = OL.half_list(s["x"], s["y"], s["z"], L)
    fx, fy, fz, o = OL.forces_list(s, L, f0, cnt, lst)
    assert _rel_force_err(md.forces(0), (fx, fy, fz)) < 1e-12
    assert pv[0] == pytest.approx(o[0], rel=1e-12)
    assert pv[1] == pytest.approx(o[1], rel=1e-12)


@pytest.mark.parametrize("n", [300, 500, 640, 874])
def test_small_box_forces_match_brute_force(gpu, n):
    """Boxes below N = 875 have fewer than 3 list cells per axis, where the
    reference's two-cell grid aliases its stencil (MDcode/Parallel.c:72,
    cell_list+forces.c:37-49) and the build visits every cell once: forces
    against the oracle's O(N^2) minimum-image sum, FP64 to 1e-12 and FP32
    to 1e-5, and the FP64 energy and virial."""
    s, L = lattice_state(n, jitter=0.06, seed=n + 3)
    assert L > 2 * 2.85  # the list radius stays below half the box
    fx, fy, fz, o = OL.forces_brute(s, L, 1.2)
    for prec, tol in ((0, 1e-12), (1, 1e-5)):
        md = make_md(gpu, [(s, L)], f0=1.2, prec=prec)
        pv = md.calculate_forces()[0]
        err = _rel_force_err(md.forces(0), (fx, fy, fz))
        print("N=%d %s: force error %.2e" % (n, ("FP64", "FP32")[prec], err))
        assert err < tol
        if prec == 0:
            assert pv[0] == pytest.approx(o[0], rel=1e-12)
            assert pv[1] == pytest.approx(o[1], rel=1e-12)
        md.close()


@pytest.mark.parametrize("skin", [0.25, 0.5])
def test_skin_changes_lists_not_forces(gpu, skin):
    """A different Verlet skin changes list radii, cell size and rebuild
    interval, never the forces: FP64 forces still match the oracle, and a
    run of graph-replayed steps reports no list overflow."""
    s, L = lattice_state(3000, jitter=0.08, seed=44)
    gpu.set_skin(skin)
    try:
        md = make_md(gpu, [(s, L)], f0=1.0)
    finally:
        gpu.set_skin(0.35)
    md.calculate_forces()
    cnt, lst, _ = OL.half_list(s["x"], s["y"], s["z"], L)
    fx, fy, fz, o = OL.forces_list(s, L, 1.0, cnt, lst)
    assert _rel_force_err(md.forces(0), (fx, fy, fz)) < 1e-12
    md.close()
    gpu.set_skin(skin)
    try:
        md = make_relaxed_md(gpu, [(2000, 45, 6)], tau=50, prec=1)
    finally:
        gpu.set_skin(0.35)
    md.step(400, graph=True)
    md.sync()  # raises on a list overflow or a runaway coordinate
    assert md.rebuilds(0) > 0


def test_forces_mixed_batch_match_oracle(gpu):
    # replicas of very different N share one launch (different lanes per
    # particle and list layouts); every one must still match the oracle
    states = [lattice_state(n, jitter=0.06, seed=n + 3) for n in (1000, 5000, 1837, 12000)]
    md = make_md(gpu, states, f0=1.0)
    pv = md.calculate_forces()
    for r, (s, L) in enumerate(states):
        cnt, lst, _ = OL.half_list(s["x"], s["y"], s["z"], L)
        fx, fy, fz, o = OL.forces_list(s, L, 1.0, cnt, lst)
        assert _rel_force_err(md.forces(r), (fx, fy, fz)) < 1e-12, r
        assert pv[r][0] == pytest.approx(o[0], rel=1e-12)


def test_forces_fp32_error_bound(gpu):
    errs = []
    for seed in range(4):
        s, L = lattice_state(3000, jitter=0.06, seed=50 + seed)
        md = make_md(gpu, [(s, L)], prec=1)
        md.calculate_forces()
        cnt, lst, _ = OL.half_list(s["x"], s["y"], s["z"], L)
        fx, fy, fz, o = OL.forces_list(s, L, 1.5, cnt, lst)
        errs.append(_rel_force_err(md.forces(0), (fx, fy, fz)))
    print("FP32 force error (max over particles, per configuration):", errs)
    assert max(errs) <= 1e-5, errs


def _chain0(n, T):
    q = np.zeros(3)
    OL.load().or_nhc_masses(n, T, OL.dp(q))
    return np.concatenate([np.zeros(3), np.sqrt(T * q) * np.array([0.4, -0.3, 0.2]), q])


def test_fresh_handle_chain_masses_known_answers(gpu):
    """The chain masses a new replica starts with (InitialiseNoseHoover,
    time_evolve.c:477-482) against hand-worked values; the trajectory test
    below passes the oracle's masses in and would not see a mass error."""
    from test_oracle_md import CHAIN_MASS_KAT
    cases = [(n, T, want) for n, T, want in CHAIN_MASS_KAT if n >= 500]
    md = gpu.MD([gpu.ReplicaParams(n, 1.2, T, 1.0, -1, 200, 5 + k) for k, (n, T, _) in enumerate(cases)])
    for k, (n, T, want) in enumerate(cases):
        c = md.get_chain(k)
        assert c[6:9] == pytest.approx(want, rel=1e-14, abs=0)
        assert np.all(c[0:6] == 0.0)
    md.close()


def test_nhc_trajectory_matches_oracle(gpu):
    n, T, f0 = 1000, 0.6, 1.5
    s, L = lattice_state(n, T=T, jitter=0.02, seed=21)
    # relax on the oracle so the compared window is ordinary dynamics
    s, ch, _ = OL.nhc_steps(s, L, T, f0, 200, _chain0(n, T))
    s = {k: v for k, v in s.items() if k not in ("fx", "fy", "fz")}
    md = make_md(gpu, [(s, L)], f0=f0, T=T)
    md.set_chain(0, ch)
    md.step(60)
    g = md.get_state(0)
    o, och, oout = OL.nhc_steps(s, L, T, f0, 60, ch)
    for k in ("xu", "yu", "zu"):
        assert np.abs(g[k] - o[k]).max() < 1e-11
    for k in ("px", "py", "pz"):
        assert np.abs(g[k] - o[k]).max() < 1e-9
    assert np.allclose(md.get_chain(0), och, rtol=1e-10, atol=1e-12)
    st = md.stats(0)
    assert st["Heff"] == pytest.approx(oout[4], rel=1e-10)
    assert st["ke"] == pytest.approx(oout[2], rel=1e-10)
    assert st["pe"] == pytest.approx(oout[0], rel=1e-10)
    assert md.rebuilds(0) >= 1


@pytest.mark.parametrize("crit", [0, 1])
def test_rebuild_criterion_matches_oracle(gpu, crit):
    """Both list-rebuild criteria against the oracle's step loop
    (time_evolve.c:547-556): the safe 4 d^2 > skin^2 (crit 0) and the
    reference's 3 d^2 > skin^2 (crit 1, set with
    ac4_md_set_rebuild_criterion).  Over 300 steps at f0=2 the replica
    rebuilds exactly as often as the oracle with the same rule, and the
    trajectories agree as in the plain NHC comparison."""
    n, T, f0 = 1000, 0.6, 2.0
    s, L = lattice_state(n, T=T, jitter=0.02, seed=23)
    s, ch, _ = OL.nhc_steps(s, L, T, f0, 200, _chain0(n, T))
    s = {k: v for k, v in s.items() if k not in ("fx", "fy", "fz")}
    md = make_md(gpu, [(s, L)], f0=f0, T=T)
    md.set_chain(0, ch)
    md.set_rebuild_criterion(crit)
    r0 = md.rebuilds(0)
    md.step(300)
    g = md.get_state(0)
    o, _, oout = OL.nhc_steps(s, L, T, f0, 300, ch, safe=1 - crit)
    assert md.rebuilds(0) - r0 == int(oout[5]) >= 5
    for k in ("xu", "yu", "zu"):
        assert np.abs(g[k] - o[k]).max() < 1e-9
    md.close()


@pytest.mark.parametrize("n", [2000, 7000, 10240, 10250, 40000])
def test_shuffle_follows_reference_swap_rule(gpu, n):
    """200, 700 and 1024 active particles (the traced shuffle, one, three and
    four positions per lane), 1025 and 4000 (the serial chain): the
    reference's swap chain, value for value, on the same uniforms, for a
    shuffle forced now and for the one the next step makes."""
    s, L = lattice_state(n, seed=4)
    md = make_md(gpu, [(s, L)], tau=200)
    before = md.get_state(0)
    u = md.shuffle_uniforms(0)
    md.shuffle()
    after = md.get_state(0)
    for c, k in enumerate(("kx", "ky", "kz")):
        ref = np.array(before[k], dtype=np.int32)
        OL.load().or_shuffle(len(ref), OL.ip(ref), OL.dp(np.ascontiguousarray(u[c])))
        assert np.array_equal(after[k], ref)
        # a permutation: the same number of +1 and -1 (the sum is 0 for an
        # even active count)
        assert after[k].sum() == np.sum(before[k]) and set(after[k].tolist()) <= {-1, 1}
        assert n % 20 != 0 or after[k].sum() == 0
    assert not np.array_equal(after["kx"], before["kx"])
    # the shuffle a step makes at step 0 of tau: inside the half-kick kernel's
    # first workgroups on handles whose shuffle LDS is small (n up to 10250
    # here), as its own kernel otherwise (40000); the next uniforms
    u2 = md.shuffle_uniforms(0)
    md.step(1)
    stepped = md.get_state(0)
    for c, k in enumerate(("kx", "ky", "kz")):
        ref = np.array(after[k], dtype=np.int32)
        OL.load().or_shuffle(len(ref), OL.ip(ref), OL.dp(np.ascontiguousarray(u2[c])))
        assert np.array_equal(stepped[k], ref)
    assert not np.array_equal(stepped["kx"], after["kx"])


@pytest.mark.parametrize("prec,tpp", [(0, 4), (1, 1), (1, 8)])
def test_replica_alone_equals_in_batch(gpu, prec, tpp):
    # the force team width fixes the summation order: pinned for both handles
    with tuning(gpu, tpp):
        alone = make_relaxed_md(gpu, [(1000, 31, 77)], tau=20, prec=prec)
        mixed = make_relaxed_md(gpu, [(1500, 32, 5), (1000, 31, 77), (2600, 33, 9)], tau=20, prec=prec)
    alone.step(150)
    alone.sync()
    mixed.step(150)
    mixed.sync()
    a, m = alone.get_state(0), mixed.get_state(1)
    for k in ("xu", "yu", "zu", "px", "py", "pz", "kx"):
        assert np.array_equal(a[k], m[k]), k
    assert alone.stats(0)["Heff"] == mixed.stats(1)["Heff"]
    assert alone.rebuilds(0) == mixed.rebuilds(1) > 0


@pytest.mark.parametrize("prec", [0, 1])
def test_list_builders_give_identical_trajectories(gpu, prec):
    """Wave-per-cell, thread-per-particle and 16/32/64-lane team list
    builders write the same neighbours in the same order, so the dynamics are
    bitwise equal."""
    spec = [(1300, 61, 4), (2400, 62, 5)]
    out = []
    for mode in (0, 1, 2, 3, 4):
        with tuning(gpu, 0, mode):
            md = make_relaxed_md(gpu, spec, tau=30, prec=prec)
        md.step(200, graph=True)
        md.sync()
        out.append([md.get_state(r) for r in range(2)] + [md.rebuilds(0), md.rebuilds(1)])
        md.close()
    assert out[0][2] > 0 and out[0][3] > 0
    for o in out[1:]:
        assert o[2:] == out[0][2:]
        for r in range(2):
            for k in ("xu", "yu", "zu", "px"):
                assert np.array_equal(o[r][k], out[0][r][k]), (r, k)


@pytest.mark.parametrize("prec", [0, 1])
def test_checkpoint_restore_is_exact(gpu, prec):
    """A replica checkpointed mid-run and restored into another slot of
    another handle continues bitwise as the uninterrupted run; a checkpoint
    does not load into a slot of a different shape."""
    with tuning(gpu, 4):
        a = make_relaxed_md(gpu, [(1500, 81, 3), (2200, 82, 4)], tau=20, prec=prec)
        b = make_relaxed_md(gpu, [(2200, 82, 4)], tau=20, prec=prec)
        c = make_relaxed_md(gpu, [(2000, 83, 5)], tau=20, prec=prec)
        d = make_relaxed_md(gpu, [(2200, 82, 9)], tau=20, prec=prec)  # same shape, another seed
    a.step(150, graph=True)
    blob = a.checkpoint(1)
    a.step(170, graph=True)
    a.sync()
    b.restore(0, blob)
    b.step(170, graph=True)
    b.sync()
    ga, gb = a.get_state(1), b.get_state(0)
    for k in ("x", "xu", "yu", "zu", "px", "py", "pz", "kx"):
        assert np.array_equal(ga[k], gb[k]), k
    assert np.array_equal(a.get_chain(1), b.get_chain(0))
    sa, sb = a.stats(1), b.stats(0)
    assert sa["Heff"] == sb["Heff"] and sa["steps"] == sb["steps"] == 320
    assert a.rebuilds(1) == b.rebuilds(0) > 0
    with pytest.raises(gpu.AC4Error):
        c.restore(0, blob)
    with pytest.raises(gpu.AC4Error, match="seed"):
        d.restore(0, blob)


def test_graph_replay_equals_eager(gpu):
    """Graph replays (chunks of 25 captured steps, then one-step graphs for
    the rest: 57 = 2 x 25 + 7, 43 = 25 + 18) equal eager launches."""
    spec = [(1200, 41, 1), (2000, 42, 2)]
    e = make_relaxed_md(gpu, spec, tau=25, prec=1)
    g = make_relaxed_md(gpu, spec, tau=25, prec=1)
    e.step(100)
    e.sync()
    g.step(57, graph=True)
    g.sync()
    g.step(43, graph=True)
    g.sync()
    for r in range(2):
        se, sg = e.get_state(r), g.get_state(r)
        for k in ("xu", "px", "kz"):
            assert np.array_equal(se[k], sg[k])
        assert e.stats(r)["steps"] == g.stats(r)["steps"] == 100


@pytest.mark.parametrize("prec", [0, 1])
def test_launch_groups_equal_one_handle(gpu, prec):
    """Replicas split into launch groups (one handle and stream each, graph
    replays interleaved by ac4_md_step_graph_multi) follow the same
    trajectories as the same replicas stepped eagerly in one handle, with
    the force team width pinned for all handles."""
    spec = [(1300, 71, 3), (2100, 72, 4), (1000, 73, 5), (1700, 74, 6)]
    with tuning(gpu, 1):
        one = make_relaxed_md(gpu, spec, tau=20, prec=prec)
        ga = make_relaxed_md(gpu, [spec[0], spec[3]], tau=20, prec=prec)
        gb = make_relaxed_md(gpu, [spec[1], spec[2]], tau=20, prec=prec)
    one.step(137)
    one.sync()
    gpu.step_multi([ga, gb], 137)
    ga.sync()
    gb.sync()
    for r, (h, q) in enumerate([(ga, 0), (gb, 0), (gb, 1), (ga, 1)]):
        a, b = one.get_state(r), h.get_state(q)
        for k in ("xu", "yu", "zu", "px", "py", "pz", "kx"):
            assert np.array_equal(a[k], b[k]), (r, k)
        assert one.stats(r)["Heff"] == h.stats(q)["Heff"]
        assert one.rebuilds(r) == h.rebuilds(q) > 0
    for m in (one, ga, gb):
        m.close()


def test_stats_consistent_with_state(gpu):
    md = make_relaxed_md(gpu, [(1000, 51, 3)], tau=50, prec=0)
    md.step(30)
    md.sync()
    st, g = md.stats(0), md.get_state(0)
    ke = 0.5 * (g["px"] ** 2 + g["py"] ** 2 + g["pz"] ** 2).sum()
    assert st["ke"] == pytest.approx(ke, rel=1e-12)
    assert st["drift"] == pytest.approx(g["px"].sum(), abs=1e-9)
    assert st["steps"] == 30


def _run_server(ac4, caps, jobs, chunk=100, overlap=None, budget=None, fail=None, groups=0):
    """Runs the jobs to completion, one tick per time unit; fail = (job
    index, tick) raises an error flag on that job's replica after the tick."""
    if budget is not None:
        srv = ac4.Server(precision=ac4.FP32, chunk=chunk, relax=200, budget=budget, n_range=caps, overlap=overlap)
    else:
        srv = ac4.Server(caps, precision=ac4.FP32, chunk=chunk, relax=200, overlap=overlap, groups=groups)
    ids = [srv.submit(*j) for j in jobs]
    t = 0
    while srv.pending() and t < 10000:
        srv.tick(float(t))
        if fail and t == fail[1]:
            srv.raise_error(ids[fail[0]], 2)
        t += 1
    return srv, ids


@pytest.mark.parametrize("groups", [1, 3])
def test_server_job_alone_equals_in_mix(gpu, groups):
    """Jobs in a mix (slots in one launch group, or in three groups stepped
    together) give bitwise the results they give alone in a one-group
    server of the same slots."""
    caps = [2048, 2048, 4096, 8192]
    # (n, T, f0, tau_p, steps, seed, t_arrive in ticks)
    jobs = [(1500, 0.5, 1.0, 1.0, 600, 11, 0), (3000, 0.6, 0.5, 0.1, 900, 12, 0), (1800, 0.45, 2.0, 10.0, 500, 13, 1),
            (7000, 0.5, 1.5, 1.0, 400, 14, 2), (1200, 0.8, 0.0, 1.0, 700, 15, 3), (2000, 0.5, 1.0, 0.5, 300, 16, 5)]
    mix, ids = _run_server(gpu, caps, jobs, groups=groups)
    assert mix.groups() == groups
    st = [mix.job(i) for i in ids]
    assert all(s["state"] == 2 for s in st)
    assert st[5]["t_start"] > 0  # the third 2048-class job had to wait for a free slot
    for k in (0, 2, 3):
        alone, aid = _run_server(gpu, caps, [jobs[k][:6] + (0,)], groups=1)
        a, m = alone.result(aid[0]), mix.result(ids[k])
        for u in range(3):
            assert np.array_equal(a[u], m[u]), (k, u)
        alone.close()
    assert mix.particle_steps() > 0


@pytest.mark.parametrize("groups", [1, 3])
def test_server_failed_job_is_isolated(gpu, groups):
    """A job whose replica raises an error flag (set from the host, as a
    kernel sets it on a list overflow or a runaway coordinate) fails alone:
    its slot is cleared and serves a later job, and every other job finishes
    bitwise as when it runs alone (in a one-group server; the mix in one or
    three launch groups)."""
    caps = [2048, 2048, 4096]
    ov = (200, 4, 8, 0.3)
    jobs = [(1500, 0.5, 1.0, 1.0, 800, 31, 0), (1800, 0.6, 0.5, 0.1, 800, 32, 0), (3000, 0.5, 2.0, 10.0, 600, 33, 0),
            (1900, 0.7, 1.5, 1.0, 500, 34, 6)]
    mix, ids = _run_server(gpu, caps, jobs, overlap=ov, fail=(1, 3), groups=groups)
    assert mix.groups() == groups
    st = [mix.job(i) for i in ids]
    assert st[1]["state"] == 3 and mix.job_error(ids[1]) == 2
    assert st[1]["steps_done"] == 400  # four chunks ran, the fifth saw the flag
    with pytest.raises(gpu.AC4Error):
        mix.result(ids[1])
    assert [st[k]["state"] for k in (0, 2, 3)] == [2, 2, 2]
    assert st[3]["t_start"] >= 6
    for k in (0, 2, 3):
        alone, aid = _run_server(gpu, caps, [jobs[k][:6] + (0,)], overlap=ov, groups=1)
        a, m = alone.result(aid[0]), mix.result(ids[k])
        for u in range(3):
            assert np.array_equal(a[u], m[u]), (k, u)
        for u, (va, vm) in enumerate(zip(alone.job_overlap(aid[0]), mix.job_overlap(ids[k]))):
            assert np.array_equal(va, vm), (k, u)
        alone.close()
    mix.close()


def test_server_cfg4_stream(gpu):
    """The serving configuration's request distributions (N logU[500, 20000],
    T U[0.45, 0.8], f0 U[0, 2], tau_p logU[0.1, 10]; steps cut to 300-1200 and
    a job below N = 875 added) on slots planned from an HBM budget, with every
    job's Q(t) kept on the device: all jobs finish, and the smallest and the
    largest job give bitwise the results (final x_u, Q sums) they give alone."""
    rng = np.random.default_rng(404)
    jobs = []
    for k in range(10):
        n = int(np.exp(rng.uniform(np.log(500), np.log(20000))))
        jobs.append((n, float(rng.uniform(0.45, 0.8)), float(rng.uniform(0.0, 2.0)),
                     float(np.exp(rng.uniform(np.log(0.1), np.log(10.0)))), int(rng.integers(300, 1201)), 500 + k,
                     float(k // 2)))
    jobs.append((600, 0.6, 1.2, 1.0, 700, 599, 1.0))
    ov = (500, 6, 12, 0.3)
    budget = 160e6
    caps = gpu.server_plan(budget, 500, 20000, 6, 12)
    assert sum(gpu.slot_bytes(int(c), 6, 12) for c in caps) <= budget
    assert min(caps) >= 512 and max(caps) >= 20000
    mix, ids = _run_server(gpu, (500, 20000), jobs, overlap=ov, budget=budget)
    assert np.array_equal(mix.slots(), caps)
    assert all(mix.job(i)["state"] == 2 for i in ids)
    order = np.argsort([j[0] for j in jobs])
    for k in (int(order[0]), int(order[-1])):
        alone, aid = _run_server(gpu, (500, 20000), [jobs[k][:6] + (0.0,)], overlap=ov, budget=budget)
        a, m = alone.result(aid[0]), mix.result(ids[k])
        for u in range(3):
            assert np.array_equal(a[u], m[u]), (k, u)
        qa, qm = alone.job_overlap(aid[0]), mix.job_overlap(ids[k])
        for u in range(3):
            assert np.array_equal(qa[u], qm[u]), (k, u)
        assert qm[2][1] > 0 and 0.0 < qm[0][1] / qm[2][1] <= 1.0  # Q at the first lag
        alone.close()
    mix.close()


def test_server_preempted_job_resumes_exactly(gpu):
    """A job taken off the GPU mid-run (exact checkpoint in host memory, its
    Q(t) origins and sums included), whose slot then serves another job,
    finishes bitwise as when it runs uninterrupted."""
    caps = [2048, 4096]
    ov = (200, 4, 6, 0.3)
    a_job = (1500, 0.5, 1.0, 1.0, 700, 21, 0)
    alone, aid = _run_server(gpu, caps, [a_job], overlap=ov)
    ref = alone.result(aid[0])
    ref_q = alone.job_overlap(aid[0])
    alone.close()
    srv = gpu.Server(caps, precision=gpu.FP32, chunk=100, relax=200, overlap=ov, groups=2)  # one slot a group
    assert srv.groups() == 2
    ia = srv.submit(*a_job)
    srv.tick(0.0)
    srv.tick(1.0)
    srv.tick(2.0)  # A has run 300 steps (relaxation ramp + chain)
    srv.preempt(ia)
    assert srv.job(ia)["state"] == 0
    ib = srv.submit(1800, 0.6, 0.5, 0.1, 500, 22, 3.0)  # takes the freed 2048 slot
    t = 3.0
    while srv.pending() and t < 200:
        srv.tick(t)
        t += 1.0
    assert srv.job(ia)["state"] == 2 and srv.job(ib)["state"] == 2
    assert srv.job(ib)["t_start"] < srv.job(ia)["t_done"]
    got = srv.result(ia)
    for u in range(3):
        assert np.array_equal(got[u], ref[u]), u
    for u, (a, b) in enumerate(zip(srv.job_overlap(ia), ref_q)):
        assert np.array_equal(a, b), u
    srv.close()


def test_server_preempted_job_resumes_in_another_group(gpu):
    """A job preempted in one launch group and readmitted into a slot of the
    other (its old slot taken meanwhile) finishes bitwise as when it runs
    uninterrupted in a one-group server of the same slots."""
    caps = [2048, 2048]
    ov = (200, 4, 6, 0.3)
    a_job = (1500, 0.5, 1.0, 1.0, 700, 21, 0)
    alone, aid = _run_server(gpu, caps, [a_job], overlap=ov, groups=1)
    ref, ref_q = alone.result(aid[0]), alone.job_overlap(aid[0])
    alone.close()
    srv = gpu.Server(caps, precision=gpu.FP32, chunk=100, relax=200, overlap=ov, groups=2)
    ia = srv.submit(*a_job)
    for t in (0.0, 1.0, 2.0):
        srv.tick(t)
    first = gpu.lib.ac4_server_job_slot(srv.h, ia)
    srv.preempt(ia)
    ib = srv.submit(1800, 0.6, 0.5, 0.1, 500, 22, 3.0)  # takes the freed slot
    srv.tick(3.0)
    assert gpu.lib.ac4_server_job_slot(srv.h, ib) == first
    second = gpu.lib.ac4_server_job_slot(srv.h, ia)
    assert second >= 0 and second != first  # readmitted into the other group's slot
    t = 4.0
    while srv.pending() and t < 200:
        srv.tick(t)
        t += 1.0
    assert srv.job(ia)["state"] == 2 and srv.job(ib)["state"] == 2
    got = srv.result(ia)
    for u in range(3):
        assert np.array_equal(got[u], ref[u]), u
    for u, (a, b) in enumerate(zip(srv.job_overlap(ia), ref_q)):
        assert np.array_equal(a, b), u
    srv.close()


def test_insitu_overlap_accumulators(gpu):
    specs = [(1000, 71, 1), (1600, 72, 2)]
    md = make_relaxed_md(gpu, specs, tau=20, prec=1)
    se, oe, no, nl = 10, 40, 6, 12
    md.overlap_enable(se, oe, no, nl, 0.3)
    frames = {}
    md.overlap_sample()  # step 0: first origin
    frames[0] = [md.get_state(r) for r in range(2)]
    for k in range(1, 13):
        md.step(se, graph=True)
        md.overlap_sample()
        frames[k * se] = [md.get_state(r) for r in range(2)]
    md.sync()
    for r, (n, _, _) in enumerate(specs):
        L = (n / 1.2) ** 0.33333333
        q, q2, c = md.overlap(r)
        wq, wq2, wc = np.zeros(nl + 1), np.zeros(nl + 1), np.zeros(nl + 1, dtype=np.int32)
        for step in sorted(frames):
            for s0 in sorted(frames):
                if s0 % oe or s0 >= step or step - s0 > nl * se or (step - s0) // oe >= no + (step - s0) // oe:
                    continue
                a, b = frames[step][r], frames[s0][r]
                d = [a[k] - b[k] for k in ("xu", "yu", "zu")]
                d = [np.where(v >= 0.5, v - 1.0, np.where(v < -0.5, v + 1.0, v)) for v in d]
                r2 = L * L * ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
                qq = float((r2 < 0.3 * 0.3).sum()) / n
                lag = (step - s0) // se
                wq[lag] += qq
                wq2[lag] += qq * qq
                wc[lag] += 1
        assert np.array_equal(c, wc)
        assert np.array_equal(q, wq) and np.array_equal(q2, wq2)


def test_fix_drift_removes_mean_momentum(gpu):
    """ac4_md_fix_drift (fixDrift, MDcode/initialise.c:104-138): the mean
    momentum, summed in reference particle order and divided by N, is
    subtracted from every particle exactly; the run then continues."""
    md = make_relaxed_md(gpu, [(1200, 91, 3)], tau=20, prec=0)
    md.step(40)
    md.sync()
    g = md.get_state(0)
    s0 = relaxed_state(gpu, 1200, 91)[0]
    drift = np.array([0.3, -0.2, 0.1])
    p = {k: g[k] + d for k, d in zip(("px", "py", "pz"), drift)}
    md.set_state(0, g["x"], g["y"], g["z"], g["xu"], g["yu"], g["zu"], p["px"], p["py"], p["pz"], s0["type"],
                 s0["act_tag"], g["kx"], g["ky"], g["kz"])
    md.init_chain(0, seed=91)
    md.step(10)
    md.sync()
    before = md.get_state(0)
    m = md.fix_drift(0)
    after = md.get_state(0)
    for a, k in enumerate(("px", "py", "pz")):
        assert m[a] == np.cumsum(before[k])[-1] / 1200  # sequential sum, as the reference's loop
        assert np.array_equal(after[k], before[k] - m[a])
        assert abs(after[k].sum()) < 1e-12
    assert np.all(np.abs(m - drift) < 0.05)
    md.step(20)
    md.sync()
    assert 0.3 < md.stats(0)["T"] < 1.2
    md.close()


def test_format_frame_on_device_matches_host(gpu):
    """ac4_md_format_frame (print_data, MDcode/print.c:4-16, formatted on the
    device) is byte for byte the host formatter's text of the same frame: two
    replicas after steps (slots permuted by the cell sort), then crafted values
    next to every rounding edge of %.5lf, and values the device leaves to the
    host (|v| >= 1e12, inf, nan)."""
    md = make_relaxed_md(gpu, [(1200, 91, 3), (1000, 31, 4)], tau=20, prec=0)
    md.step(60)
    md.sync()
    s0 = [relaxed_state(gpu, n, sd)[0] for n, sd in ((1200, 91), (1000, 31))]
    for r in range(2):
        g = md.get_state(r)
        want = gpu.format_frame(g["xu"], g["yu"], g["zu"], s0[r]["type"])
        assert md.format_frame(r) == want
    n = 1200
    g = md.get_state(0)
    rng = np.random.default_rng(5)
    k = rng.integers(-10**9, 10**9, size=n).astype(float)
    edge = (k + 0.5) / 1e5  # nearest doubles to the rounding boundaries
    vals = np.concatenate([edge, np.nextafter(edge, np.inf), np.nextafter(edge, -np.inf)])
    special = np.array([0.0, -0.0, 5e-324, -5e-324, -4e-6, -6e-6, 4.9999999e-6, 5.0000001e-6, 1e-5, 0.99999,
                        0.999995, 9.99999999999e11, -9.99999999999e11, 2.0 ** -60, 123456.654321])
    xu = vals[:n].copy()
    xu[:len(special)] = special
    yu, zu = vals[n:2 * n].copy(), vals[2 * n:3 * n].copy()
    md.set_state(0, g["x"], g["y"], g["z"], xu, yu, zu, g["px"], g["py"], g["pz"], s0[0]["type"], s0[0]["act_tag"],
                 g["kx"], g["ky"], g["kz"])
    assert md.format_frame(0) == gpu.format_frame(xu, yu, zu, s0[0]["type"])
    for bad in (1e12, -3.5e13, np.inf, np.nan):
        xb = xu.copy()
        xb[7] = bad
        md.set_state(0, g["x"], g["y"], g["z"], xb, yu, zu, g["px"], g["py"], g["pz"], s0[0]["type"],
                     s0[0]["act_tag"], g["kx"], g["ky"], g["kz"])
        assert md.format_frame(0) == gpu.format_frame(xb, yu, zu, s0[0]["type"])
    md.close()


def test_parallel_shuffle_follows_reference_swap_rule(gpu):
    """Active sets above 16384 take the parallel shuffle (per-block source
    maps of the swap sequence, applied in block order): the directions equal
    the oracle's serial swap chain (MDcode/active.c:58-75) on the same
    uniforms, for a shuffle forced now and for the one a step makes."""
    n = 200000  # 20000 active particles
    s, L = lattice_state(n, seed=11, jitter=0.0)
    md = make_md(gpu, [(s, L)], tau=1)
    for it in range(2):
        before = md.get_state(0)
        u = md.shuffle_uniforms(0)
        if it == 0:
            md.shuffle()
        else:
            md.step(1)  # tau = 1: the step shuffles first
            md.sync()
        after = md.get_state(0)
        for c, k in enumerate(("kx", "ky", "kz")):
            ref = np.array(before[k], dtype=np.int32)
            OL.load().or_shuffle(len(ref), OL.ip(ref), OL.dp(np.ascontiguousarray(u[c])))
            assert np.array_equal(after[k], ref), (it, k)
            assert after[k].sum() == 0
        assert not np.array_equal(after["kx"], before["kx"])
    md.close()


def test_serial_and_parallel_shuffles_in_one_handle(gpu):
    """A handle whose replicas take both shuffle paths (200 and 17000 active
    particles): each replica's directions equal the oracle's swap chain."""
    states = [lattice_state(n, seed=12 + k, jitter=0.0) for k, n in enumerate((2000, 170000))]
    md = make_md(gpu, states, tau=200)
    before = [md.get_state(r) for r in range(2)]
    u = [md.shuffle_uniforms(r) for r in range(2)]
    md.shuffle()
    for r in range(2):
        after = md.get_state(r)
        for c, k in enumerate(("kx", "ky", "kz")):
            ref = np.array(before[r][k], dtype=np.int32)
            OL.load().or_shuffle(len(ref), OL.ip(ref), OL.dp(np.ascontiguousarray(u[r][c])))
            assert np.array_equal(after[k], ref), (r, k)
    md.close()


def test_list_teams_on_a_large_handle(gpu):
    """Above 4e5 particles the 32-lane teams leave the copy back to k_commit:
    their run equals the 64-lane teams' bit for bit (same lists, same
    order)."""
    out = []
    for mode in (3, 4):
        with tuning(gpu, 0, mode):
            s, L = lattice_state(420000, seed=13, jitter=0.0)
            md = make_md(gpu, [(s, L)], tau=50, prec=1)
            md.relax(100, 1.5, 1.0)
            md.init_chain(0, seed=13)
        md.step(60, graph=True)
        md.sync()
        g = md.get_state(0)
        out.append((g, md.rebuilds(0)))
        md.close()
    assert out[0][1] == out[1][1] > 0
    for k in ("xu", "yu", "zu", "px"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
