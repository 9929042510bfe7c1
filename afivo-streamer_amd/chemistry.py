"""Reaction lists in the reference's input format, compiled for the device
chemistry (afgpu_chem_*): host-side set-up, run once per simulation.

read_network follows read_reactions (src/m_chemistry.f90:857-1142),
parse_reaction (:1156-1278), table_from_file (src/m_table_data.f90:136-220)
and to_simple_ascii (src/m_chemistry.f90:1359-1400) for gases of constant
density: gas species and "M" fold into the rate factor, species are
numbered in order of first appearance, field tables are interpolated onto
the uniform lookup grid as LT_get_spaced_data does (linear, clamped).
Reaction groups ("@" lines) are expanded as read_reactions expands them
(src/m_chemistry.f90:918-975).
"""
import ctypes as C
import re

import numpy as np

import afgpu as A

RATE_TYPES = {
    "c1": (A.RATE_CONSTANT, 1),
    "c1*(Td-c2)": (A.RATE_LINEAR, 2),
    "c1*exp(-(c2/(c3+Td))**2)": (A.RATE_EXP_V1, 3),
    "c1*exp(-(Td/c2)**2)": (A.RATE_EXP_V2, 2),
    "c1*(300/Te)**c2": (A.RATE_K1, 2),
    "(c1*(kB_eV*Te+c2)**2-c3)*c4": (A.RATE_K3, 4),
    "c1*(Tg/300)**c2*exp(-c3/Tg)": (A.RATE_K4, 3),
    "c1*exp(-c2/Tg)": (A.RATE_K5, 2),
    "c1*Tg**c2": (A.RATE_K6, 2),
    "c1*(Tg/c2)**c3": (A.RATE_K7, 3),
    "c1*(300/Tg)**c2": (A.RATE_K8, 2),
    "c1*exp(-c2*Tg)": (A.RATE_K9, 2),
    "10**(c1+c2*(Tg-300))": (A.RATE_K10, 2),
    "c1*(300/Tg)**c2*exp(-c3/Tg)": (A.RATE_K11, 3),
    "c1*Tg**c2*exp(-c3/Tg)": (A.RATE_K12, 3),
    "c1*exp(-(c2/(c3+Td))**c4)": (A.RATE_K13, 4),
    "c1*exp(-(Td/c2)**c3)": (A.RATE_K14, 3),
    "c1*exp(-(c2/(kb*(Tg+Td/c3)))**c4)": (A.RATE_K15, 4),
    "c1*(300/Te)**c2*exp(-c3/Tg)*exp(c4*(Te-Tg)/(Te*Tg))": (A.RATE_K16, 4),
}


def simple_ascii(name):
    """to_simple_ascii: the output name and the charge of a species."""
    out, charge, brackets = "", 0, False
    for ch in name:
        if ch == "(":
            brackets = True
            out += "_"
        elif ch == ")":
            brackets = False
        elif ch == "*":
            out += "_star"
        elif ch == "+":
            charge += 0 if brackets else 1
            out += "_plus"
        elif ch == "-":
            charge -= 0 if brackets else 1
            out += "_min"
        elif ch == "^":
            out += "_hat"
        elif ch == "'":
            out += "p"
        else:
            out += ch
    if out in ("e", "stoch_e"):
        charge = -1
    return out, charge


def _block(lines, start):
    """Lines after a header line and its line of dashes, up to the next
    dashes."""
    i = start + 1
    if not lines[i].strip().startswith("-----"):
        raise ValueError("%s not followed by -----" % lines[start].strip())
    out = []
    for ln in lines[i + 1:]:
        s = ln.strip()
        if s.startswith("-----"):
            return out
        out.append(s)
    raise ValueError("no closing dashes after %s" % lines[start].strip())


def table_from_file(lines, name):
    """table_from_file: the (x, y) data below the line equal to name, after
    its FACTOR:/COMMENT: lines and dashes, up to the next dashes."""
    for k, ln in enumerate(lines):
        if ln.rstrip("\n").rstrip() == name:
            break
    else:
        raise ValueError("table %r not found" % name)
    factor = 1.0
    i = k + 1
    while True:
        s = lines[i].strip()
        i += 1
        if s.startswith("-----"):
            break
        if s.startswith("FACTOR:"):
            factor = float(s[7:])
        elif not s.startswith("COMMENT:"):
            raise ValueError("unknown statement %r in table %r" % (s, name))
    rows = []
    while True:
        s = lines[i].strip()
        i += 1
        if s.startswith("-----"):
            break
        if s == "" or s.startswith("#"):
            continue
        rows.append([float(v) for v in s.replace(",", " ").split()[:2]])
    xy = np.asarray(rows)
    return xy[:, 0], xy[:, 1] * factor


def spaced_data(x, y, grid):
    """LT_get_spaced_data / LT_lin_interp_list: linear, clamped at both ends."""
    out = np.empty(len(grid))
    for n, g in enumerate(grid):
        if g <= x[0]:
            out[n] = y[0]
        elif g >= x[-1]:
            out[n] = y[-1]
        else:
            i = int(np.searchsorted(x, g, side="left"))
            t = (g - x[i - 1]) / (x[i] - x[i - 1])
            out[n] = (1 - t) * y[i - 1] + t * y[i]
    return out


class Network:
    """A compiled reaction list: species (input names, output names,
    charges), reactions (afgpu_reaction_t array), field-table columns."""

    def __init__(self):
        self.species, self.names, self.charges = [], [], []
        self.reactions, self.descriptions, self.tables = [], [], []


def lt_grid(x_min, x_max, n):
    """The rows of a linearly spaced lookup table, as LT_create forms them
    (get_x, src/lookup_table_fortran/m_lookup_table.f90:303-327):
    x_min + ((i - 1) * (1 / (n - 1))) * (x_max - x_min)."""
    tmp = 1.0 / (n - 1)
    return x_min + (np.arange(n) * tmp) * (x_max - x_min)


def electrons_first(net):
    """The device path takes species 0 as the electrons (the reference finds
    them by name, "e"): move "e" to the front, keeping the other species in
    their order of first appearance, and renumber the reactions."""
    if "e" not in net.species or net.species[0] == "e":
        return
    order = ["e"] + [sp for sp in net.species if sp != "e"]
    new_ix = {old: order.index(sp) for old, sp in enumerate(net.species)}
    for r in net.reactions:
        for q in range(r.n_in):
            r.ix_in[q] = new_ix[r.ix_in[q]]
        for q in range(r.n_out):
            r.ix_out[q] = new_ix[r.ix_out[q]]
    net.species = order


MAX_GROUPS = 10  # max_groups, src/m_chemistry.f90:881


def parse_group(line, groups):
    """An "@name = a, b, ..." line (src/m_chemistry.f90:920-940): the fields
    between "=" and "," delimiters, the first the name, the others the
    members; every group of one reaction has the same number of members."""
    fields = [f.strip() for f in re.split(r"[=,]", line) if f.strip()]
    if len(groups) + 1 > MAX_GROUPS:
        raise ValueError("Too many groups")
    if groups and len(fields) - 1 != len(groups[0][1]):
        raise ValueError("Groups for a reaction should have the same size: %s" % line)
    return fields[0], fields[1:]


def string_replace(text, old, new):
    """Every occurrence of old replaced, scanning again from the start after
    each replacement (string_replace, src/m_chemistry.f90:1334-1356)."""
    while True:
        i = text.find(old)
        if i < 0:
            return text
        text = text[:i] + new + text[i + len(old):]


def expand_groups(entries, groups):
    """The reaction read before the "@" lines becomes one reaction per group
    member: member k of every group replaces the group's name in the
    reaction, its rate type and its data, groups in the order given
    (src/m_chemistry.f90:942-975).  The reference does this when it reads
    the next reaction line, so groups after the list's last reaction are not
    applied.  Reactions that come out the same in text and data are
    refused, as there."""
    if not entries:
        raise ValueError("reaction groups before any reaction")
    template = entries.pop()
    new = []
    for k in range(len(groups[0][1])):
        reaction, how, data, unit = template
        for name, members in groups:
            reaction = string_replace(reaction, name, members[k])
            how = string_replace(how, name, members[k])
            data = string_replace(data, name, members[k])
        new.append((reaction, how, data, unit))
    for e in new:
        if sum(1 for x in new if x[0] == e[0]) > 1 and sum(1 for x in new if x[2] == e[2]) > 1:
            raise ValueError("Groups lead to duplicate reactions: %s" % e[0])
    entries.extend(new)


def read_network(path, gas_fractions, gas_number_density, table_x_min, table_x_max, table_size):
    """The reaction list of an input file for a gas of constant density
    with the given component fractions (e.g. {"N2": 0.8, "O2": 0.2})."""
    lines = open(path).read().splitlines()
    ignored = []
    for k, ln in enumerate(lines):
        if ln.strip() == "ignored_species":
            ignored = _block(lines, k)
    starts = [k for k, ln in enumerate(lines) if ln.strip() == "reaction_list"]
    if not starts:
        raise ValueError("no reaction_list in %s" % path)
    entries = []
    groups = []  # (name, members) of the "@" lines after a reaction
    for s in _block(lines, starts[0]):
        if s.startswith("#") or s == "":
            continue
        if s.startswith("@"):
            groups.append(parse_group(s, groups))
            continue
        if groups:
            expand_groups(entries, groups)
            groups = []
        f = [t.strip() for t in s.split(",")]
        if not 3 <= len(f) <= 4:
            raise ValueError("invalid chemistry syntax: %s" % s)
        entries.append((f[0], f[1], f[2], f[3] if len(f) > 3 else "m"))
    net = Network()
    grid = lt_grid(table_x_min, table_x_max, table_size)
    for text, how, data, unit in entries:
        r = A.ReactionT()
        rfactor, n_species_in, left = 1.0, 0, True
        ins, outs = [], {}
        keep = True
        for comp in text.split():
            if comp == "+":
                continue
            if comp == "->":
                left = False
                continue
            mult = 1
            if comp[0] in "123456789":
                mult, comp = int(comp[0]), comp[1:]
            if left:
                n_species_in += mult
            if comp in gas_fractions:
                if left:
                    rfactor *= gas_fractions[comp] * gas_number_density
                continue
            if comp == "M":
                if left:
                    rfactor *= gas_number_density
                continue
            if comp in ignored:
                if left:
                    keep = False
                    break
                continue
            if comp not in net.species:
                net.species.append(comp)
            ix = net.species.index(comp)
            if left:
                ins += [ix] * mult
            else:
                outs[ix] = outs.get(ix, 0) + mult
        if not keep:
            continue
        if not ins:
            raise ValueError("no input species: %s" % text)
        if how == "field_table":
            x, y = table_from_file(lines, data)
            r.rate_type = A.RATE_TABULATED_FIELD
            r.lookup_col = len(net.tables)
            net.tables.append(spaced_data(x, y, grid))
        elif how in RATE_TYPES:
            r.rate_type, n_coeff = RATE_TYPES[how]
            vals = [float(v) for v in re.split(r"[\s,]+", data.strip())[:n_coeff]]
            for q, v in enumerate(vals):
                r.rate_data[q] = v
        else:
            raise ValueError("unknown rate type %r" % how)
        if unit == "cm":
            rfactor *= 1e-6 ** (n_species_in - 1)
        elif unit != "m":
            raise ValueError("invalid length unit %r" % unit)
        r.rate_factor = rfactor
        r.n_in, r.n_out = len(ins), len(outs)
        for q, ix in enumerate(ins):
            r.ix_in[q] = ix
        for q, (ix, m) in enumerate(outs.items()):
            r.ix_out[q] = ix
            r.multiplicity_out[q] = m
        net.reactions.append(r)
        net.descriptions.append(text)
    electrons_first(net)
    for name in net.species:
        simple, charge = simple_ascii(name)
        net.names.append(simple)
        net.charges.append(charge)
    net.table_x_min, net.table_x_max, net.table_size = table_x_min, table_x_max, table_size
    net.fld_cols = np.ascontiguousarray(np.array(net.tables).reshape(len(net.tables), table_size))
    return net


def reaction_array(reactions):
    arr = (A.ReactionT * max(1, len(reactions)))()
    for k, r in enumerate(reactions):
        arr[k] = r
    return arr


class Chem:
    """The device chemistry (kind 'gpu': arrays in device memory, such as
    afgpu.DeviceArray, anything with data_ptr()) or the oracle's (kind
    'oracle': numpy arrays)."""

    def __init__(self, kind, n_species, reactions, table_x_min, table_x_max, fld_cols, energy_col=None,
                 gas_temperature=300.0):
        self.kind = kind
        self.lib = A.load_gpu() if kind == "gpu" else A.load_oracle()
        self.pre = "afgpu_chem_" if kind == "gpu" else "afcpu_chem_"
        self.n_species, self.n_reactions = n_species, len(reactions)
        fld = np.ascontiguousarray(fld_cols, np.float64)
        n_cols = fld.shape[0] if fld.size else 0
        table_n = fld.shape[1] if fld.size else (len(energy_col) if energy_col is not None else 2)
        en = None if energy_col is None else np.ascontiguousarray(energy_col, np.float64)
        self.h = C.c_void_p()
        rc = getattr(self.lib, self.pre + "create")(
            n_species, len(reactions), reaction_array(reactions), table_n, table_x_min, table_x_max, n_cols,
            A._dp(fld) if n_cols else None, A._dp(en) if en is not None else None, gas_temperature,
            C.byref(self.h))
        if rc != A.OK:
            raise A.AfError("%screate failed (%d): %s" % (self.pre, rc, A.last_error(self.kind)))

    def _ptr(self, a):
        if a is None:
            return None
        if self.kind == "gpu":
            return C.c_void_p(a.data_ptr())
        return a.ctypes.data_as(C.c_void_p)

    def _call(self, name, *args):
        rc = getattr(self.lib, self.pre + name)(self.h, *args)
        if rc != A.OK:
            raise A.AfError("%s%s failed (%d): %s" % (self.pre, name, rc, A.last_error(self.kind)))

    def rates(self, fields, out):
        self._call("rates", len(fields), self._ptr(fields), self._ptr(out))

    def derivatives(self, dens, rates, derivs):
        self._call("derivatives", derivs.shape[-1], self._ptr(dens), self._ptr(rates), self._ptr(derivs))

    def source(self, fields, dens, derivs):
        self._call("source", len(fields), self._ptr(fields), self._ptr(dens), self._ptr(derivs))

    def close(self):
        if getattr(self, "h", None):
            getattr(self.lib, self.pre + "destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
