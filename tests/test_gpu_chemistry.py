"""General chemistry on the device against the oracle: the reference's air
network (tests/golden/air_chemistry_v1.txt: 8 species, 25 reactions, field
tables and analytic rates) and a network with every rate type, on random
fields and densities (cell counts not a multiple of the workgroup).

Rates built from the table, constants and products only are equal to the
oracle's bit for bit.  Those with exp or pow use the device library's,
which may differ from the host's in the last bit; exp(-y) turns a last-bit
difference of y into about |y| ulps of the rate, so rates are compared to
4e-15 (1 + |y|) relative, y = ln(rate / (c0 c1)), and derivatives to the
sum of their terms' bounds.  The fused source equals the two-pass derivatives bit for bit
on the device.  Arrays live in device memory (afgpu.DeviceArray).
"""
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
GOLD = os.path.join(ROOT, "tests", "golden")
N_GAS = 2.414e25
AIR = {"N2": 0.8, "O2": 0.2}

ALL_TYPES = """reaction_list
-----
e + N2 -> e + e + N2+,field_table,T1
e + O2 -> O-,c1*(Td-c2),1e-18 20.0
O- + N2 -> e,c1*exp(-(c2/(c3+Td))**2),1.16e-18 48.9 11.0
O- + O2 + M -> O3-,c1*exp(-(Td/c2)**2),1.1e-42 65.0
e + N2+ ->,c1*(300/Te)**c2,2e-13 0.7
e + O3- -> O- + e,(c1*(kB_eV*Te+c2)**2-c3)*c4,1e-16 0.1 1e-19 2.0
N2+ + O- ->,c1*(Tg/300)**c2*exp(-c3/Tg),1e-13 -0.5 20.0
N2+ + O3- ->,c1*exp(-c2/Tg),2e-13 100.0
O- + O3- -> O- + O-,c1*Tg**c2,1e-20 1.5
N2+ + N2 + M -> N4+,c1*(Tg/c2)**c3,5e-41 300.0 -1.64
N4+ + O- ->,c1*(300/Tg)**c2,1e-13 0.5
N4+ + O2 -> N2+,c1*exp(-c2*Tg),2.5e-16 0.001
N4+ + O3- ->,10**(c1+c2*(Tg-300)),-13.0 0.002
N4+ + e -> N2+ + e,c1*(300/Tg)**c2*exp(-c3/Tg),1e-16 0.3 50.0
O3- + N2 -> O- + N2,c1*Tg**c2*exp(-c3/Tg),1e-18 0.5 400.0
O- + M -> e,c1*exp(-(c2/(c3+Td))**c4),1e-18 50.0 10.0 1.5
O3- + M -> e,c1*exp(-(Td/c2)**c3),1e-20 80.0 2.5
N2+ + M -> N4+,c1*exp(-(c2/(kb*(Tg+Td/c3)))**c4),1e-42 4e-21 0.18 1.0
e + N4+ ->,c1*(300/Te)**c2*exp(-c3/Tg)*exp(c4*(Te-Tg)/(Te*Tg)),3e-13 0.5 10.0 20.0
-----

T1
COMMENT: test table
-----
0.0 0.0
100.0 1e-17
400.0 5e-16
1000.0 3e-15
-----
"""


def network(CH, tmp_path, which, size):
    if which == "air":
        path = os.path.join(GOLD, "air_chemistry_v1.txt")
    else:
        path = str(tmp_path / "all_types.txt")
        open(path, "w").write(ALL_TYPES)
    return CH.read_network(path, AIR, N_GAS, 0.0, 1000.0, size)


def inputs(ns, n, seed):
    rng = np.random.default_rng(seed)
    fields = rng.uniform(0.0, 1100.0, n)  # |E| in Td: on and above the table
    fields[:5] = [0.0, 1000.0, 500.0, 1e-3, 1e4]
    dens = 10.0 ** rng.uniform(10, 20, (ns, n))
    return fields, dens


@pytest.mark.parametrize("which, n", [("air", 100003), ("all_types", 70001)])
def test_chemistry_equals_oracle(A, CH, tmp_path, which, n):
    net = network(CH, tmp_path, which, 1000)
    ns, nr = len(net.species), len(net.reactions)
    energy = np.linspace(0.2, 25.0, 1000)
    mk = lambda kind: CH.Chem(kind, ns, net.reactions, 0.0, 1000.0, net.fld_cols, energy, 310.0)
    g, o = mk("gpu"), mk("oracle")
    fields, dens = inputs(ns, n, 7)
    ro, do = np.zeros((nr, n)), np.zeros((ns, n))
    o.rates(fields, ro)
    r_ref = ro.copy()
    o.derivatives(dens, ro, do)
    tf = A.DeviceArray(None, fields)
    td = A.DeviceArray(None, dens)
    tr = A.DeviceArray((nr, n))
    tdv = A.DeviceArray((ns, n))
    g.rates(tf, tr)
    rg = tr.numpy()
    exact = [k for k, r in enumerate(net.reactions) if r.rate_type in (1, 2, 3)]
    assert np.array_equal(rg[exact], r_ref[exact]), which
    # exp(-y) carries the last-bit difference of y (from pow or division)
    # as a relative difference of about |y| ulps: the bound grows with the
    # exponent |ln(rate / (c0 c1))|
    tol = np.zeros((nr, n))
    for k, r in enumerate(net.reactions):
        base = abs(r.rate_factor * (1.0 if r.rate_type == A.RATE_K10 else r.rate_data[0]))
        with np.errstate(divide="ignore", invalid="ignore"):
            y = np.abs(np.log(np.abs(r_ref[k]) / base))
        tol[k] = 4e-15 * (1 + np.where(np.isfinite(y), y, 0.0))
    bad = np.abs(rg - r_ref) > tol * np.abs(r_ref) + 1e-300
    assert not bad.any(), (which, np.argwhere(bad)[:5], rg[bad][:5], r_ref[bad][:5])
    g.derivatives(td, tr, tdv)
    dg = tdv.numpy()
    # bound: the terms of species s are its reactions' rates (times
    # multiplicity), each within its rate's bound
    scale = np.zeros((ns, n))
    for k, r in enumerate(net.reactions):
        t = np.abs(ro[k]) * (tol[k] + 1e-14)
        for q in range(r.n_in):
            scale[r.ix_in[q]] += t
        for q in range(r.n_out):
            scale[r.ix_out[q]] += t * r.multiplicity_out[q]
    assert np.all(np.abs(dg - do) <= scale + 1e-300), which
    assert np.all(np.abs(tr.numpy() - ro) <= (tol + 1e-14) * np.abs(ro) + 1e-300)
    ts = A.DeviceArray((ns, n))
    g.source(tf, td, ts)
    assert np.array_equal(ts.numpy(), dg), which


@pytest.mark.parametrize("ns, nr", [(5, 15), (10, 30), (20, 60), (60, 60), (60, 180), (100, 300)])
def test_chemistry_many_species(A, CH, ns, nr):
    """Random networks of nr constant-rate reactions over ns species (every
    workgroup shape: 256 or 64 cells, densities and reaction records in LDS
    or not): bit for bit equal to the oracle (no exp or pow)."""
    rng = np.random.default_rng(ns + nr)
    reactions = []
    for _ in range(nr):
        r = A.ReactionT()
        r.rate_type = A.RATE_CONSTANT
        r.rate_factor = 10.0 ** rng.uniform(-20, -14)
        r.rate_data[0] = rng.uniform(0.5, 2.0)
        r.n_in = int(rng.integers(1, 4))
        r.n_out = int(rng.integers(0, 4))
        for q in range(r.n_in):
            r.ix_in[q] = int(rng.integers(0, ns))
        for q in range(r.n_out):
            r.ix_out[q] = int(rng.integers(0, ns))
            r.multiplicity_out[q] = int(rng.integers(1, 3))
        reactions.append(r)
    cols = np.zeros((0, 16))
    g = CH.Chem("gpu", ns, reactions, 0.0, 1000.0, cols, None, 300.0)
    o = CH.Chem("oracle", ns, reactions, 0.0, 1000.0, cols, None, 300.0)
    n = 4099
    fields = rng.uniform(0.0, 500.0, n)
    dens = 10.0 ** rng.uniform(10, 18, (ns, n))
    do = np.zeros((ns, n))
    o.source(fields, dens, do)
    out = A.DeviceArray((ns, n))
    g.source(A.DeviceArray(None, fields), A.DeviceArray(None, dens), out)
    assert np.array_equal(out.numpy(), do)
