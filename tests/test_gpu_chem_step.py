"""The time step with a reaction list (afgpu_set_chemistry).

- The standard model written as a reaction list (the two tabulated
  reactions chemistry_initialize builds without a list,
  src/m_chemistry.f90:226-258, on the same table) gives the standard
  model's steps bit for bit, on the device and in the oracle.
- A richer list over the same three species (detachment with an exp-type
  rate, electron-ion and ion-ion recombination, a gas-temperature rate):
  device against oracle.  Its exp and pow may differ in the last bit between
  device and host, so the states are compared to 1e-11 of each variable's
  largest magnitude after three steps (they agree to about 1e-15 per rate).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def standard_list(A):
    r1, r2 = A.ReactionT(), A.ReactionT()
    for r, col in ((r1, 0), (r2, 1)):
        r.rate_type, r.lookup_col, r.rate_factor, r.n_in = A.RATE_TABULATED_FIELD, col, 1.0, 1
        r.ix_in[0] = 0
    r1.n_out, r2.n_out = 2, 1
    r1.ix_out[0], r1.multiplicity_out[0], r1.ix_out[1], r1.multiplicity_out[1] = 0, 2, 1, 1
    r2.ix_out[0], r2.multiplicity_out[0] = 2, 1
    return [r1, r2]


def rich_list(A, N):
    out = standard_list(A)

    def add(t, c0, coef, ins, outs):
        r = A.ReactionT()
        r.rate_type, r.rate_factor, r.n_in, r.n_out = t, c0, len(ins), len(outs)
        for q, v in enumerate(coef):
            r.rate_data[q] = v
        for q, s in enumerate(ins):
            r.ix_in[q] = s
        for q, (s, m) in enumerate(outs):
            r.ix_out[q], r.multiplicity_out[q] = s, m
        out.append(r)
    add(A.RATE_EXP_V1, N, [1.24e-17, 179.0, 8.8], [2], [(0, 1)])      # M- + M -> e
    add(A.RATE_CONSTANT, 1.0, [1e-13], [0, 1], [])                    # e + M+ ->
    add(A.RATE_CONSTANT, 1.0, [2e-13], [1, 2], [])                    # M+ + M- ->
    add(A.RATE_K8, 1.0, [3e-14, 0.7], [0, 1], [])                     # e + M+ -> (Tg)
    return out


def chem_for(S, CH, case, kind, reactions, n_species=3):
    td = case.tdata
    n, x0 = td["n"], td["chem_x_min"]
    # the last grid point as the tables' set-up forms it (the field range of
    # the transport data, in Td), so that LT_create's inv_fac is the table's
    E = S.analytic_transport()[0]
    x_max_td = E[-1] * 1e21 / case.params.gas_number_density
    tmp = 1.0 / (n - 1)
    xl = (n - 1) * tmp
    x_last = 0.0 + xl * (x_max_td - 0.0)
    assert (n - 1) / (x_last - x0) == td["chem_inv_fac"]
    cols = np.asarray(td["chem"]).reshape(2, n)
    return CH.Chem(kind, n_species, reactions, x0, x_last, cols, None, 300.0)


def cases(S):
    return {
        "cyl2d": S.small_amr(2, coord=2, uniform_lvl=3, max_lvl=5, frac=0.3, seed=3),
        "xyz3d_8": S.small_amr(3, nc=8, uniform_lvl=2, max_lvl=4, frac=0.3, seed=11),
    }


def run(A, S, case, kind, chem, steps=3, seed_neg=False, charges=None, species=None):
    if kind == "gpu":
        p = A.Path("gpu", case.D, case.nc, case.coord, A.PREC_FP64)
    else:
        p = A.Path("oracle", case.D, case.nc, case.coord)
    S.apply(p, case)
    if seed_neg:  # negative ions (a tenth of the electrons), neutral overall
        p.put_cc(A.CC_NEG_ION, 0.1 * np.asarray(case.ne0))
        p.put_cc(A.CC_POS_ION, 1.1 * np.asarray(case.ne0))
    if chem is not None:
        p.set_chemistry(chem, charges)
    for k, dens in (species or {}).items():
        p.put_species(k, dens)
    p.field_compute(0, have_guess=False)
    infos = [p.step() for _ in range(steps)]
    return p, infos


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8"])
def test_standard_model_as_a_list_equals_standard_steps(A, S, CH, name):
    case = cases(S)[name]
    for kind in ("gpu", "oracle"):
        ref, ri = run(A, S, case, kind, None)
        chem = chem_for(S, CH, case, kind, standard_list(A))
        p, pi = run(A, S, case, kind, chem)
        assert pi == ri, (name, kind)
        for iv in range(A.N_CC):
            assert np.array_equal(p.get_cc(iv), ref.get_cc(iv)), (name, kind, iv)


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8"])
def test_reaction_list_steps_equal_oracle(A, S, CH, name):
    case = cases(S)[name]
    N = case.params.gas_number_density
    g, gi = run(A, S, case, "gpu", chem_for(S, CH, case, "gpu", rich_list(A, N)), seed_neg=True)
    o, oi = run(A, S, case, "oracle", chem_for(S, CH, case, "oracle", rich_list(A, N)), seed_neg=True)
    for a, b in zip(gi, oi):
        assert a["n_rejected"] == b["n_rejected"] and a["n_vcycles"] == b["n_vcycles"]
        assert a["dt_next"] == pytest.approx(b["dt_next"], rel=1e-11)
    for iv in (A.CC_ELECTRON, A.CC_POS_ION, A.CC_NEG_ION, A.CC_PHI, A.CC_EFIELD):
        x, y = g.get_cc(iv), o.get_cc(iv)
        assert np.max(np.abs(x - y)) <= 1e-11 * np.max(np.abs(y)), (name, iv)
    # the list changed the state: electrons and ions recombine, negative
    # ions detach and recombine
    std, _ = run(A, S, case, "oracle", None, seed_neg=True)
    for iv in (A.CC_ELECTRON, A.CC_POS_ION, A.CC_NEG_ION):
        assert not np.array_equal(std.get_cc(iv), o.get_cc(iv)), (name, iv)


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8"])
def test_general_species_path_reproduces_standard_steps(A, S, CH, name):
    """The standard model as a list over four species, the fourth neutral
    and inert, takes the general path (pools beyond the three, the rhs over
    the charged species, k_update_n) and still gives the standard steps bit
    for bit, on the device and in the oracle."""
    case = cases(S)[name]
    for kind in ("gpu", "oracle"):
        ref, ri = run(A, S, case, kind, None)
        chem = chem_for(S, CH, case, kind, standard_list(A), n_species=4)
        p, pi = run(A, S, case, kind, chem, charges=[-1, 1, -1, 0])
        assert pi == ri, (name, kind)
        for iv in range(A.N_CC):
            assert np.array_equal(p.get_cc(iv), ref.get_cc(iv)), (name, kind, iv)
        assert not p.get_species(3).any()


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8"])
def test_air_network_steps_equal_oracle(A, S, CH, name):
    """The reference's air network (8 species, 25 reactions, read from
    tests/golden/air_chemistry_v1.txt) in the time step: device against
    oracle after three steps, every species to 1e-11 of its largest
    magnitude (exp-type rates differ in the last bit)."""
    import os
    from conftest import ROOT
    case = cases(S)[name]
    N = case.params.gas_number_density
    E = S.analytic_transport()[0]
    x_max = E[-1] * 1e21 / N
    net = CH.read_network(os.path.join(ROOT, "tests", "golden", "air_chemistry_v1.txt"), {"N2": 0.8, "O2": 0.2}, N,
                          0.0, x_max, 1000)
    ns = len(net.species)
    ne0 = np.asarray(case.ne0)
    seed = {1: 0.8 * ne0, 2: 0.2 * ne0}  # N2+ and O2+ neutralise the electrons
    out = {}
    for kind in ("gpu", "oracle"):
        chem = CH.Chem(kind, ns, net.reactions, 0.0, x_max, net.fld_cols, None, 300.0)
        p, infos = run(A, S, case, kind, chem, charges=net.charges, species=seed)
        out[kind] = (p, infos, chem)
    (g, gi, _), (o, oi, _) = out["gpu"], out["oracle"]
    for a, b in zip(gi, oi):
        assert a["n_rejected"] == b["n_rejected"] and a["n_vcycles"] == b["n_vcycles"]
    for k in range(ns):
        x, y = g.get_species(k), o.get_species(k)
        assert np.max(np.abs(x - y)) <= 1e-11 * max(np.max(np.abs(y)), 1e-300), (name, k, net.names[k])
    assert o.get_species(ns - 1).any()  # O4+ formed


def test_air_network_regression_log(A, S, CH, tmp_path):
    """The regression log of a run with the air network: its header is the
    reference's own test_cyl_chem_rtest.log header (8 species), and its
    statistics equal the oracle's (maxima exactly, sums to 1e-12 plus the
    run's 1e-11 state difference)."""
    import json
    import os
    from conftest import ROOT
    case = cases(S)["cyl2d"]
    N = case.params.gas_number_density
    E = S.analytic_transport()[0]
    x_max = E[-1] * 1e21 / N
    net = CH.read_network(os.path.join(ROOT, "tests", "golden", "air_chemistry_v1.txt"), {"N2": 0.8, "O2": 0.2}, N,
                          0.0, x_max, 1000)
    ns = len(net.species)
    ne0 = np.asarray(case.ne0)
    ivs = [A.VAR_SPECIES(k) for k in range(ns)]
    vals, keep = {}, []
    fn = str(tmp_path / "chem_rtest.log")
    for kind in ("gpu", "oracle"):
        chem = CH.Chem(kind, ns, net.reactions, 0.0, x_max, net.fld_cols, None, 300.0)
        keep.append(chem)
        p, _ = run(A, S, case, kind, chem, charges=net.charges, species={1: 0.8 * ne0, 2: 0.2 * ne0})
        vals[kind] = np.concatenate([p.regression_values(ivs[:4]), p.regression_values(ivs[4:])], axis=1)
        if kind == "gpu":
            p.write_regression_log(fn, 0, ivs[:8], net.names)
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "regression_logs.json")))
    assert open(fn).read().splitlines()[0] == ref["programs/standard_2d/tests/test_cyl_chem_rtest.log"]["header"]
    g, o = vals["gpu"], vals["oracle"]
    assert np.allclose(g[..., :2], o[..., :2], rtol=1e-10, atol=0)
    assert np.allclose(g[..., 2], o[..., 2], rtol=1e-11, atol=0)


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8"])
def test_air_network_new_boxes_equal_oracle(A, S, CH, name):
    """Restriction into coarsened boxes and prolongation into new boxes
    (auto_restrict / auto_prolong) cover every species of the list: from
    the same state, every species equals the oracle's bit for bit; a step on
    the new tree then matches to 1e-11."""
    import os
    from conftest import ROOT
    case = cases(S)[name]
    N = case.params.gas_number_density
    E = S.analytic_transport()[0]
    x_max = E[-1] * 1e21 / N
    net = CH.read_network(os.path.join(ROOT, "tests", "golden", "air_chemistry_v1.txt"), {"N2": 0.8, "O2": 0.2}, N,
                          0.0, x_max, 1000)
    ns = len(net.species)
    t0 = case.extra["export"]
    n0 = t0["n_boxes"]
    rng = np.random.default_rng(1)
    dens = {k: 10.0 ** rng.uniform(12, 16, (n0, (case.nc + 2) ** case.D)) for k in range(1, ns)}
    paths, keep = {}, []
    for kind in ("gpu", "oracle"):
        chem = CH.Chem(kind, ns, net.reactions, 0.0, x_max, net.fld_cols, None, 300.0)
        keep.append(chem)
        if kind == "gpu":
            p = A.Path("gpu", case.D, case.nc, case.coord, A.PREC_FP64)
        else:
            p = A.Path("oracle", case.D, case.nc, case.coord)
        S.apply(p, case)
        p.set_chemistry(chem, net.charges)
        for k, d in dens.items():
            p.put_species(k, d)
        p.field_compute(0, have_guess=False)
        paths[kind] = p
    g, o = paths["gpu"], paths["oracle"]
    nch = 2 ** case.D
    ch = t0["children"].reshape(n0, nch)
    derefine = [b + 1 for b in range(n0) if ch[b, 0] > 0 and (ch[ch[b] - 1, 0] == 0).all()]
    leaves = [b + 1 for b in range(n0) if ch[b, 0] == 0
              and all(v != 0 for v in t0["neighbors"][b * 2 * case.D:(b + 1) * 2 * case.D])]
    pick = sorted(rng.choice(leaves, size=max(1, len(leaves) // 4), replace=False))
    for p in (g, o):
        p.restrict_boxes(derefine)
    for k in range(ns):
        assert np.array_equal(g.get_species(k), o.get_species(k)), (name, "restrict", k)
    case.tree.refine(np.asarray(pick, np.int32))
    t1 = case.tree.export()
    new = np.arange(n0 + 1, t1["n_boxes"] + 1, dtype=np.int32)
    for p in (g, o):
        p.sync_tree(t1)
        p.prolong_boxes(new)
    for k in range(ns):
        assert np.array_equal(g.get_species(k), o.get_species(k)), (name, "prolong", k)
    for p in (g, o):
        p.field_compute(0, have_guess=True)
        p.step()
    for k in range(ns):
        x, y = g.get_species(k), o.get_species(k)
        assert np.max(np.abs(x - y)) <= 1e-11 * np.max(np.abs(y)), (name, "step", k)


def test_air_network_forest_equals_single_runs(A, S, CH):
    """A forest of runs with the air network: every run's steps and all
    eight species equal the same run alone, bit for bit."""
    import os
    from conftest import ROOT
    from test_gpu_ensemble import STEP_KEYS
    ens = S.ensemble(level_list=[4, 3, 5], num_vcycles=20, rel_residual=1e-10)
    case0 = ens.cases[0]
    N = case0.params.gas_number_density
    E = S.analytic_transport()[0]
    x_max = E[-1] * 1e21 / N
    net = CH.read_network(os.path.join(ROOT, "tests", "golden", "air_chemistry_v1.txt"), {"N2": 0.8, "O2": 0.2}, N,
                          0.0, x_max, 1000)
    ns = len(net.species)
    chem = CH.Chem("gpu", ns, net.reactions, 0.0, x_max, net.fld_cols, None, 300.0)
    g = A.Path("gpu", 2, 8, A.CYL, A.PREC_FP64)
    S.apply_ensemble(g, ens)
    g.set_chemistry(chem, net.charges)
    ne = g.get_cc(A.CC_ELECTRON)
    g.put_species(1, 0.8 * ne)
    g.put_species(2, 0.2 * ne)
    singles = []
    for case in ens.cases:
        s = A.Path("gpu", 2, 8, A.CYL, A.PREC_FP64)
        S.apply(s, case)
        s.set_chemistry(chem, net.charges)
        ne = s.get_cc(A.CC_ELECTRON)
        s.put_species(1, 0.8 * ne)
        s.put_species(2, 0.2 * ne)
        singles.append(s)
    g.field_compute_runs(0, have_guess=False)
    for s in singles:
        s.field_compute(0, have_guess=False)
    for it in range(3):
        infos = g.step_ensemble()
        for r, s in enumerate(singles):
            ref = s.step()
            assert {k: infos[r][k] for k in STEP_KEYS} == {k: ref[k] for k in STEP_KEYS}, (it, r)
            ids = ens.run_ids(r)
            for k in range(ns):
                assert np.array_equal(g.get_species(k, ids=ids), s.get_species(k)), (it, r, net.names[k])


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8"])
def test_autotune_with_air_network_leaves_every_species(A, S, CH, name):
    """Tuning the work decomposition with a reaction list set runs trial
    steps from a snapshot: every species pool (those beyond the three
    standard pools included) is restored, so the following steps equal an
    untuned handle's bit for bit, and the tuned steps equal the oracle's to
    1e-11 (exp-type rates differ in the last bit)."""
    import os
    from conftest import ROOT
    case = cases(S)[name]
    N = case.params.gas_number_density
    E = S.analytic_transport()[0]
    x_max = E[-1] * 1e21 / N
    net = CH.read_network(os.path.join(ROOT, "tests", "golden", "air_chemistry_v1.txt"), {"N2": 0.8, "O2": 0.2}, N,
                          0.0, x_max, 1000)
    ns = len(net.species)
    ne0 = np.asarray(case.ne0)
    seed = {1: 0.8 * ne0, 2: 0.2 * ne0}
    paths, keep = {}, []
    for kind in ("tuned", "plain", "oracle"):
        chem = CH.Chem("oracle" if kind == "oracle" else "gpu", ns, net.reactions, 0.0, x_max, net.fld_cols, None,
                       300.0)
        keep.append(chem)
        p, _ = run(A, S, case, "oracle" if kind == "oracle" else "gpu", chem, steps=1, charges=net.charges,
                   species=seed)
        paths[kind] = p
    t, u, o = paths["tuned"], paths["plain"], paths["oracle"]
    before = [t.get_species(k) for k in range(ns)]
    t.autotune()
    for k in range(ns):
        assert np.array_equal(t.get_species(k), before[k]), (name, "restored", net.names[k])
    for it in range(2):
        a, b, c = t.step(), u.step(), o.step()
        assert a == b, (name, it)
        assert a["n_vcycles"] == c["n_vcycles"] and a["n_rejected"] == c["n_rejected"], (name, it)
    for k in range(ns):
        x, y, z = t.get_species(k), u.get_species(k), o.get_species(k)
        assert np.array_equal(x, y), (name, net.names[k])
        assert np.max(np.abs(x - z)) <= 1e-11 * max(np.max(np.abs(z)), 1e-300), (name, net.names[k])


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8"])
def test_box_io_of_species_and_faces_equals_oracle(A, S, CH, name):
    """Checkpoint box data (afgpu_get_boxes / afgpu_put_boxes,
    af_write_tree's order) over the air network's species and both face
    variables: on the device each block equals the variable's own transfer,
    and device and oracle write and read the same data bit for bit."""
    import os
    from conftest import ROOT
    case = cases(S)[name]
    N = case.params.gas_number_density
    E = S.analytic_transport()[0]
    x_max = E[-1] * 1e21 / N
    net = CH.read_network(os.path.join(ROOT, "tests", "golden", "air_chemistry_v1.txt"), {"N2": 0.8, "O2": 0.2}, N,
                          0.0, x_max, 1000)
    ns = len(net.species)
    ne0 = np.asarray(case.ne0)
    ivs = [A.VAR_SPECIES(k) for k in (0, 3, 6, 7)] + [A.CC_PHI, A.VAR_FACE(A.FC_FLUX), A.VAR_FACE(A.FC_FIELD)]
    hs, keep = {}, []
    for kind in ("gpu", "oracle"):
        chem = CH.Chem(kind, ns, net.reactions, 0.0, x_max, net.fld_cols, None, 300.0)
        keep.append(chem)
        hs[kind], _ = run(A, S, case, kind, chem, steps=1, charges=net.charges, species={1: 0.8 * ne0, 2: 0.2 * ne0})
    g, o = hs["gpu"], hs["oracle"]
    n = g.n_boxes
    rng = np.random.default_rng(5)
    ids = np.arange(1, n + 1, dtype=np.int32)
    rng.shuffle(ids)
    ids = ids[: max(2, n // 2)]
    got = g.get_boxes(ivs, ids)
    blocks = [g.get_species(k, ids=ids) for k in (0, 3, 6, 7)] + [g.get_cc(A.CC_PHI, ids), g.get_fc(A.FC_FLUX, ids),
                                                                 g.get_fc(A.FC_FIELD, ids)]
    assert np.array_equal(got, np.concatenate(blocks, axis=1))
    data = rng.normal(size=got.shape)
    for p in (g, o):
        p.put_boxes(ivs, data, ids)
    assert np.array_equal(g.get_boxes(ivs, ids), o.get_boxes(ivs, ids))
    assert np.array_equal(g.get_boxes(ivs, ids), data)
    assert np.array_equal(g.get_fc(A.FC_FIELD, ids), o.get_fc(A.FC_FIELD, ids))
