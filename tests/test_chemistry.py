"""General chemistry (CPU): the reaction-list reader and the oracle's
get_rates / get_derivatives (src/m_chemistry.f90:672-804).

Pins:
- the reference's chemistry input (tests/golden/air_chemistry_v1.txt) reads
  into the species, in the order and with the output names, of the header
  of the regression log its rtest wrote (test_cyl_chem_rtest.log), and
  every reaction conserves charge (check_charge_conservation);
- the derivatives conserve charge for any densities: sum_s q_s dn_s/dt = 0
  up to rounding, and mass action in closed form for single reactions
  (gas species and "M" folded into the rate, multiplicities, cm units);
- each rate type against closed forms at special arguments and against the
  types it reduces to (k13 with c4 = 2 is exp_v1, k11 with c3 = 0 is k8,
  ...), so a transposed coefficient or a wrong sign fails.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden")
N_GAS = 2.414e25
AIR = {"N2": 0.8, "O2": 0.2}


def air(CH, size=200):
    return CH.read_network(os.path.join(GOLD, "air_chemistry_v1.txt"), AIR, N_GAS, 0.0, 1000.0, size)


def test_air_species_order_and_names_match_rtest_log(CH):
    net = air(CH)
    logs = json.load(open(os.path.join(GOLD, "regression_logs.json")))
    header = logs["programs/standard_2d/tests/test_cyl_chem_rtest.log"]["header"].split()
    names = [h[4:-1] for h in header if h.startswith("sum(") and not h.endswith("^2)")]
    assert net.names == names
    assert len(net.reactions) == 25
    assert net.charges == [-1, 1, 1, -1, -1, -1, 1, 1]
    q = np.array(net.charges)
    for r in net.reactions:
        qin = sum(q[r.ix_in[k]] for k in range(r.n_in))
        qout = sum(q[r.ix_out[k]] * r.multiplicity_out[k] for k in range(r.n_out))
        assert qin == qout


def test_air_derivatives_conserve_charge(A, CH):
    net = air(CH)
    ns, nr = len(net.species), len(net.reactions)
    en = np.linspace(0.5, 20.0, net.table_size)  # (no Te types in this network)
    c = CH.Chem("oracle", ns, net.reactions, 0.0, 1000.0, net.fld_cols, en)
    rng = np.random.default_rng(0)
    n = 4000
    fields = rng.uniform(0.0, 1200.0, n)
    dens = 10.0 ** rng.uniform(12, 20, (ns, n))
    rates = np.zeros((nr, n))
    derivs = np.zeros((ns, n))
    c.rates(fields, rates)
    assert np.all(rates >= 0)
    c.derivatives(dens, rates, derivs)
    q = np.array(net.charges, float)[:, None]
    scale = np.abs(q * derivs).sum(axis=0) + np.abs(rates).sum(axis=0)
    assert np.all(np.abs((q * derivs).sum(axis=0)) <= 1e-13 * scale)
    d2 = np.zeros_like(derivs)
    c.source(fields, dens, d2)
    assert np.array_equal(d2, derivs)


def single(CH, tmp_path, text, ns_expect=None):
    path = tmp_path / "chem.txt"
    path.write_text("reaction_list\n-----\n%s\n-----\n" % text)
    net = CH.read_network(str(path), AIR, N_GAS, 0.0, 1000.0, 50)
    if ns_expect is not None:
        assert len(net.species) == ns_expect
    return net


def run(CH, net, fields, dens, tg=300.0, energy=None):
    c = CH.Chem("oracle", len(net.species), net.reactions, 0.0, 1000.0,
                net.fld_cols if len(net.tables) else np.zeros((0, 50)),
                energy if energy is not None else np.linspace(1.0, 10.0, 50), tg)
    n = len(fields)
    rates = np.zeros((len(net.reactions), n))
    derivs = np.zeros((len(net.species), n))
    c.rates(np.asarray(fields, float), rates)
    r0 = rates.copy()
    c.derivatives(np.asarray(dens, float), rates, derivs)
    return r0, derivs


def test_mass_action_closed_forms(CH, tmp_path):
    k = 3.0e-16
    # attachment on a gas species: dn_e = -k [O2] n_e
    net = single(CH, tmp_path, "e + O2 -> O-,c1,%g" % k, 2)
    ne = np.array([1e15, 3e17])
    _, d = run(CH, net, [100.0, 200.0], [ne, [0.0, 0.0]])
    assert np.allclose(d[0], -k * 0.2 * N_GAS * ne, rtol=1e-15)
    assert np.allclose(d[1], k * 0.2 * N_GAS * ne, rtol=1e-15)
    # ionisation: two electrons out (multiplicity 2), one in
    net = single(CH, tmp_path, "e + N2 -> e + e + N2+,c1,%g" % k, 2)
    _, d = run(CH, net, [100.0, 200.0], [ne, [0.0, 0.0]])
    assert np.allclose(d[0], k * 0.8 * N_GAS * ne, rtol=1e-15)
    assert np.allclose(d[1], k * 0.8 * N_GAS * ne, rtol=1e-15)
    # three-body with M and a cm^6/s rate: 1e-12 per extra input species
    net = single(CH, tmp_path, "A+ + B- + M ->,c1,%g,cm" % k, 2)
    na, nb = np.array([2e16, 5e17]), np.array([7e16, 1e15])
    _, d = run(CH, net, [0.0, 0.0], [na, nb])
    want = -k * 1e-12 * N_GAS * na * nb
    assert np.allclose(d[0], want, rtol=1e-14) and np.allclose(d[1], want, rtol=1e-14)
    # a species twice on the left: 2 A -> B removes A twice
    net = single(CH, tmp_path, "A + A -> B,c1,%g" % k, 2)
    _, d = run(CH, net, [0.0], [[4e16], [0.0]])
    assert d[0, 0] == pytest.approx(-2 * k * 4e16 * 4e16, rel=1e-15)
    assert d[1, 0] == pytest.approx(k * 4e16 * 4e16, rel=1e-15)


def rate_of(CH, tmp_path, how, coeffs, fields, tg=300.0, energy=None):
    net = single(CH, tmp_path, "A -> B,%s,%s" % (how, " ".join(repr(float(c)) for c in coeffs)))
    r, _ = run(CH, net, fields, [np.ones(len(fields)), np.zeros(len(fields))], tg, energy)
    return r[0]


def test_rate_types_closed_forms_and_reductions(CH, tmp_path):
    f = np.array([0.0, 37.5, 120.0, 480.0])
    R = lambda how, cs, **kw: rate_of(CH, tmp_path, how, cs, f, **kw)
    assert np.all(R("c1", [2.5e-15]) == 2.5e-15)
    lin = R("c1*(Td-c2)", [2.0, 37.5])
    assert lin[1] == 0 and lin[0] == pytest.approx(-75.0) and lin[3] == pytest.approx(2 * 442.5)
    v1 = R("c1*exp(-(c2/(c3+Td))**2)", [3.0, 47.5, 10.0])
    assert v1[1] == pytest.approx(3.0 * math.exp(-1.0), rel=1e-15)  # c2/(c3+Td) = 1
    assert np.allclose(R("c1*exp(-(c2/(c3+Td))**c4)", [3.0, 47.5, 10.0, 2.0]), v1, rtol=1e-14)
    v2 = R("c1*exp(-(Td/c2)**2)", [4.0, 120.0])
    assert v2[0] == 4.0 and v2[2] == pytest.approx(4.0 * math.exp(-1.0), rel=1e-15)
    assert np.allclose(R("c1*exp(-(Td/c2)**c3)", [4.0, 120.0, 2.0]), v2, rtol=1e-14)
    for tg in (250.0, 300.0, 411.0):
        k8 = R("c1*(300/Tg)**c2", [5.0, 1.7], tg=tg)[0]
        assert k8 == pytest.approx(5.0 * (300 / tg) ** 1.7, rel=1e-15)
        assert R("c1*(300/Tg)**c2*exp(-c3/Tg)", [5.0, 1.7, 0.0], tg=tg)[0] == pytest.approx(k8, rel=1e-15)
        k4 = R("c1*(Tg/300)**c2*exp(-c3/Tg)", [5.0, -1.7, 0.0], tg=tg)[0]
        assert k4 == pytest.approx(k8, rel=1e-14)  # (Tg/300)^-c2 = (300/Tg)^c2
        assert R("c1*(Tg/c2)**c3", [5.0, 300.0, -1.7], tg=tg)[0] == pytest.approx(k8, rel=1e-14)
        k6 = R("c1*Tg**c2", [2.0, 0.5], tg=tg)[0]
        assert k6 == pytest.approx(2.0 * math.sqrt(tg), rel=1e-15)
        assert R("c1*Tg**c2*exp(-c3/Tg)", [2.0, 0.5, 0.0], tg=tg)[0] == pytest.approx(k6, rel=1e-15)
        k5 = R("c1*exp(-c2/Tg)", [7.0, tg], tg=tg)[0]
        assert k5 == pytest.approx(7.0 / math.e, rel=1e-15)
        assert R("c1*exp(-c2*Tg)", [7.0, 1 / tg], tg=tg)[0] == pytest.approx(7.0 / math.e, rel=1e-15)
        assert R("10**(c1+c2*(Tg-300))", [-15.0, 0.01], tg=tg)[0] == pytest.approx(10 ** (-15.0 + 0.01 * (tg - 300)),
                                                                                  rel=1e-14)
        # k15 with c3 -> inf (no field term): c1 exp(-(c2 / (kB Tg))^c4)
        kB = 1.380649e-23
        k15 = R("c1*exp(-(c2/(kb*(Tg+Td/c3)))**c4)", [3.0, kB * tg, 1e300, 1.5], tg=tg)[0]
        assert k15 == pytest.approx(3.0 * math.exp(-1.0), rel=1e-14)


def test_rate_types_with_electron_temperature(CH, tmp_path):
    """Te = 2 e / (3 kB) * mean energy (eV) from the energy column: with a
    constant column of 3 kB 300 / (2 e) eV, Te is 300 K everywhere."""
    kB, e = 1.380649e-23, 1.602176634e-19
    en = np.full(50, 3 * kB * 300.0 / (2 * e))
    f = np.array([0.0, 400.0])
    R = lambda how, cs: rate_of(CH, tmp_path, how, cs, f, tg=300.0, energy=en)
    assert np.allclose(R("c1*(300/Te)**c2", [6.0, 2.3]), 6.0, rtol=1e-14)
    assert np.allclose(R("c1*(300/Te)**c2*exp(-c3/Tg)*exp(c4*(Te-Tg)/(Te*Tg))", [6.0, 2.3, 0.0, 55.0]), 6.0,
                       rtol=1e-14)
    k3 = R("(c1*(kB_eV*Te+c2)**2-c3)*c4", [2.0, 1.0, 0.5, 3.0])
    x = kB / e * 300.0 + 1.0
    assert np.allclose(k3, (2.0 * x * x - 0.5) * 3.0, rtol=1e-14)


def test_tabulated_rates_interpolate_the_table(CH, tmp_path):
    """A field table on the lookup grid: exact at the data points, linear
    in between, clamped outside (LT_get_spaced_data, LT_get_col)."""
    path = tmp_path / "chem.txt"
    path.write_text("reaction_list\n-----\nA -> B,field_table,T1\n-----\n\nT1\nCOMMENT: x\n-----\n"
                    "0.0 1.0\n500.0 11.0\n1000.0 21.0\n-----\n")
    net = CH.read_network(str(path), AIR, N_GAS, 0.0, 1000.0, 101)
    f = np.array([-5.0, 0.0, 250.0, 500.0, 730.0, 1000.0, 2000.0])
    r, _ = run(CH, net, f, [np.ones(len(f)), np.zeros(len(f))])
    assert np.allclose(r[0], [1.0, 1.0, 6.0, 11.0, 15.6, 21.0, 21.0], rtol=1e-13)


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8"])
def test_oracle_standard_model_as_a_list_equals_standard_steps(A, S, CH, name):
    """The oracle's time step with the standard model written as a reaction
    list (afcpu_set_chemistry) is the standard step, bit for bit."""
    import test_gpu_chem_step as T
    case = T.cases(S)[name]
    ref, ri = T.run(A, S, case, "oracle", None)
    p, pi = T.run(A, S, case, "oracle", T.chem_for(S, CH, case, "oracle", T.standard_list(A)))
    assert pi == ri
    for iv in range(A.N_CC):
        assert np.array_equal(p.get_cc(iv), ref.get_cc(iv)), iv


@pytest.mark.parametrize("kind", ["gpu", "oracle"])
def test_invalid_networks_are_refused(A, CH, kind):
    """Rate type 7 (absent from the reference), the electron-energy table
    (needs the energy equation) and a species index out of range are
    refused before anything is allocated (no GPU use)."""
    def net(t, ix_in=0):
        r = A.ReactionT()
        r.rate_type, r.rate_factor, r.n_in = t, 1.0, 1
        r.ix_in[0] = ix_in
        r.rate_data[0] = 1.0
        return [r]
    cols = np.zeros((0, 16))
    for bad in (net(7), net(A.RATE_TABULATED_ENERGY), net(A.RATE_CONSTANT, ix_in=3)):
        with pytest.raises(A.AfError):
            CH.Chem(kind, 3, bad, 0.0, 1000.0, cols, None, 300.0)


def test_lookup_grid_equals_reference_lt_create(CH):
    """The rows of the chemistry's field table are LT_create's (get_x,
    src/lookup_table_fortran/m_lookup_table.f90:303-327), bit for bit
    against the rows the reference's own LT_create produced
    (tests/golden/ref/pointwise.npz, oracle/ref)."""
    import refvec
    d = refvec.load("pointwise")
    x = d["td.x"]
    g = CH.lt_grid(0.0, float(x[-1]), len(x))
    assert np.array_equal(g.view(np.uint64), x.view(np.uint64))
    # (the other grouping of the same product differs in the last bit)
    alt = 0.0 + np.arange(len(x)) * ((x[-1] - 0.0) / (len(x) - 1))
    assert not np.array_equal(alt, x)


def test_electrons_are_moved_first(CH, tmp_path):
    """A list whose first species is not the electron (the reference finds
    the electrons by name): "e" becomes species 0, the others keep their
    order, and every reaction's species indices follow."""
    net = single(CH, tmp_path, "O2- + M -> e + O2 + M,c1,1e-18\ne + O2 -> O2-,c1,2e-16", 2)
    assert net.species == ["e", "O2-"] and net.charges == [-1, -1]
    r0, r1 = net.reactions
    assert list(r0.ix_in[:r0.n_in]) == [1] and list(r0.ix_out[:r0.n_out]) == [0]
    assert list(r1.ix_in[:r1.n_in]) == [0] and list(r1.ix_out[:r1.n_out]) == [1]


def test_reaction_groups_expand_as_read_reactions(CH, tmp_path):
    """"@" groups (src/m_chemistry.f90:918-975): the reaction before the
    group lines becomes one reaction per member, member k of every group
    replacing the group's name everywhere in the reaction and its data
    ("@x" inside "@x+" too), when the next reaction line is read."""
    text = ("e + @x -> e + e + @x+,c1,@k\n@x = N2, O2\n@k = 1e-16, 2e-16\n"
            "e + O2 -> O2-,c1,3e-17")
    net = single(CH, tmp_path, text)
    assert net.species == ["e", "N2+", "O2+", "O2-"]
    assert [r.rate_data[0] for r in net.reactions] == [1e-16, 2e-16, 3e-17]
    assert net.reactions[0].rate_factor == 0.8 * N_GAS and net.reactions[1].rate_factor == 0.2 * N_GAS
    assert CH.string_replace("@x + @x+ -> @xy", "@x", "O2") == "O2 + O2+ -> O2y"
    # the same member twice with different data is allowed (98air_2water's H2O)
    net = single(CH, tmp_path, "e + @x -> e + e + @x+,c1,@k\n@x = N2, N2\n@k = 1e-16, 2e-16\ne -> e,c1,0")
    assert len(net.reactions) == 3
    with pytest.raises(ValueError, match="duplicate"):
        single(CH, tmp_path, "e + @x -> e + e + @x+,c1,1e-16\n@x = N2, N2\ne -> e,c1,0")
    with pytest.raises(ValueError, match="same size"):
        single(CH, tmp_path, "e + @x -> e + e + @x+,c1,@k\n@x = N2, O2\n@k = 1e-16\ne -> e,c1,0")
    # groups after the list's last reaction are not applied (the reference
    # expands on reading the next reaction line)
    net = single(CH, tmp_path, "e + O2 -> O2-,c1,3e-17\n@x = N2, O2")
    assert net.species == ["e", "O2-"] and len(net.reactions) == 1
