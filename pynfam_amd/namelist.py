"""The pnFAM input namelist: groups, defaults and a reader for the Fortran
namelist text pynfam writes (pynfam/fortran/pnfam_run.py:209-270 through
f90nml).

Groups and defaults follow exes/pnfam/pnfam_setup.f90:98-108 and :178-210; the
groups are read in the reference's order (general, ext_field, interaction,
solver, :238-241), and a group missing from the file keeps its defaults.  An
entry written without a value (`vpair_t1 = ,`, f90nml's rendering of None) keeps
the default too.
"""
from __future__ import annotations

import re

HUGE = 1.7976931348623157e308


def defaults():
    return {
        "general": {"fam_output_filename": "", "print_stdout": True, "use_fam_storage": 0,
                    "real_eqrpa": 0.0, "imag_eqrpa": 0.5},
        "ext_field": {"beta_type": "-", "operator_name": "F", "operator_k": 0, "compute_crossterms": False,
                      "two_body_current_mode": 0, "two_body_current_usep": False,
                      "two_body_current_lecs": [-3.4, 5.4, 0.0]},
        "solver": {"max_iter": 200, "broyden_history_size": 50, "convergence_epsilon": 1.0e-7,
                   "energy_shift_prot": 0.0, "energy_shift_neut": 0.0, "quench_residual_int": 1.0},
        "interaction": {"interaction_name": "NONE", "require_self_consistency": True,
                        "require_gauge_invariance": False, "force_j2_terms": False,
                        "vpair0": -HUGE, "vpair1": -HUGE, "vpair_t0": -HUGE, "vpair_t1": -HUGE,
                        "override_cs0": HUGE, "override_csr": HUGE, "override_cds": HUGE, "override_ct": HUGE,
                        "override_cf": HUGE, "override_cgs": HUGE, "override_cj": HUGE, "override_csdj": HUGE},
    }


class NamelistError(ValueError):
    pass


def _value(tok):
    t = tok.strip()
    if t == "":
        return None
    if t[0] in "'\"":
        return t[1:-1] if len(t) >= 2 and t[-1] == t[0] else t[1:]
    low = t.lower()
    if low in (".true.", "t", ".t."):
        return True
    if low in (".false.", "f", ".f."):
        return False
    try:
        return int(t)
    except ValueError:
        pass
    try:
        return float(low.replace("d", "e"))
    except ValueError:
        raise NamelistError(f"cannot read value {tok!r}")


def _split_values(rhs):
    """Comma-separated values of one entry, quotes respected."""
    out, cur, q = [], "", None
    for ch in rhs:
        if q:
            cur += ch
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            cur += ch
        elif ch == ",":
            out.append(cur)
            cur = ""
        else:
            cur += ch
    out.append(cur)
    return out


def parse(text):
    """Groups of a namelist file as {group: {name: value}} (names lower case)."""
    text = re.sub(r"!.*", "", text)
    groups = {}
    for m in re.finditer(r"&(\w+)(.*?)/", text, re.S):
        name = m.group(1).lower()
        body = m.group(2)
        entries = {}
        # split on "name =" occurrences
        parts = re.split(r"(\b\w+(?:\(\d+\))?)\s*=", body)
        for key, rhs in zip(parts[1::2], parts[2::2]):
            vals = [_value(v) for v in _split_values(rhs.strip().rstrip(","))] if rhs.strip() else [None]
            vals = [v for v in vals]
            if all(v is None for v in vals):
                continue
            entries[key.lower()] = vals[0] if len(vals) == 1 else vals
        groups[name] = entries
    return groups


def read(path):
    """Defaults overlaid with the groups of the file."""
    with open(path) as fh:
        found = parse(fh.read())
    nml = defaults()
    for g, entries in found.items():
        if g not in nml:
            continue
        for k, v in entries.items():
            if k not in nml[g]:
                raise NamelistError(f"unknown entry {k!r} in group &{g}")
            nml[g][k] = v
    return nml
