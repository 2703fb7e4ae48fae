"""Reader for the reference's Fortran namelist input files (input.dat, input_stat.dat).

The reference reads six groups from input.dat (2d_mrlbm.f90:2227-2233, 2245-2262) and
one from input_stat.dat (2d_statistics.f90:238-247).  Both shipped files end their last
group without the closing '/', which gfortran accepts; this reader accepts it too.
A malformed file raises ``NamelistError`` (the reference prints "Error while reading
input.dat file..." and stops, 2d_mrlbm.f90:2235-2238).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

INPUT_GROUPS = {
    "domainsize": ("nx", "ny", "lattice_type"),
    "cylinder": ("ncy", "l_front", "l_back", "l_top", "l_bot"),
    "numbers": ("re",),
    "parallel": ("num_threads",),
    "controls": ("uo", "rho_infty", "iplot", "max_iter", "isave", "irestart", "statsbegin",
                 "statsend", "iplotstats"),
    "logicalcontrols": ("post_process", "x_periodic", "y_periodic", "channel_with_cylinder",
                        "channel_with_square", "incomp_bc", "incomp_bulk", "vel_interp",
                        "mom_interp", "rotated_coordinate", "strong_mass_bc", "eq_mass_bc",
                        "weak_mass_bc"),
}
STAT_GROUPS = {"parameters": ("nbct", "r_cy", "uo", "rho_infty")}


class NamelistError(ValueError):
    pass


def _value(tok: str):
    t = tok.strip()
    tl = t.lower()
    if tl in (".true.", ".t.", "t", "true"):
        return True
    if tl in (".false.", ".f.", "f", "false"):
        return False
    try:
        if re.fullmatch(r"[+-]?\d+", t):
            return int(t)
        return float(tl.replace("d", "e"))
    except ValueError:
        if len(t) >= 2 and t[0] == t[-1] and t[0] in "'\"":
            return t[1:-1]
        raise NamelistError(f"cannot parse value {tok!r}")


def parse_namelists(text: str, groups: dict) -> dict:
    """Returns {group: {name: value}} for the named groups found in `text`."""
    out: dict = {}
    # drop comments
    lines = [ln.split("!", 1)[0] for ln in text.replace("\r", "").split("\n")]
    body = "\n".join(lines)
    for m in re.finditer(r"&(\w+)(.*?)(?=(/\s*(?:\n|$))|&\w+|\Z)", body, flags=re.S):
        name = m.group(1).lower()
        content = m.group(2)
        vals = {}
        for a in re.finditer(r"(\w+)\s*=\s*([^,\n]+)", content):
            key = a.group(1).lower()
            vals[key] = _value(a.group(2).strip().rstrip(","))
        out[name] = vals
    for g, names in groups.items():
        if g not in out:
            raise NamelistError(f"namelist group &{g} missing")
        for k in out[g]:
            if k not in names:
                raise NamelistError(f"unknown variable {k!r} in &{g}")
    return out


@dataclass
class RunConfig:
    """The reference's run configuration after read_input_file and the geometry
    derivation of 2d_mrlbm.f90:86-92 and 946-948."""
    nl: dict = field(default_factory=dict)

    def get(self, group, key, default=None):
        return self.nl.get(group, {}).get(key, default)

    @property
    def channel_with_cylinder(self):
        return bool(self.get("logicalcontrols", "channel_with_cylinder", True))

    @property
    def ncy(self):
        return int(self.get("cylinder", "ncy"))

    def grid(self):
        if self.channel_with_cylinder:
            ncy = self.ncy
            nfront = int(self.get("cylinder", "l_front")) * ncy
            nback = int(self.get("cylinder", "l_back")) * ncy
            ntop = int(self.get("cylinder", "l_top")) * ncy
            nbot = int(self.get("cylinder", "l_bot")) * ncy
            return nfront + ncy + nback, ntop + ncy + nbot
        return int(self.get("domainsize", "nx")), int(self.get("domainsize", "ny"))

    def centre(self):
        """(x_c, y_c, r_c), 1-based lattice coordinates (2d_mrlbm.f90:946-948)."""
        ncy = self.ncy
        nfront = int(self.get("cylinder", "l_front")) * ncy
        nbot = int(self.get("cylinder", "l_bot")) * ncy
        return float(nfront + ncy // 2 + 1), float(nbot + ncy // 2), float(ncy // 2)


def read_input(path: str) -> RunConfig:
    with open(path) as fh:
        return RunConfig(parse_namelists(fh.read(), INPUT_GROUPS))


def read_input_stat(path: str) -> dict:
    with open(path) as fh:
        return parse_namelists(fh.read(), STAT_GROUPS)["parameters"]
