"""``fv_run.dat`` reader with the semantics of the reference's FVCOM-style
scanner (``SCAN_FILE``, src/fv_utilit.F90:223-570) and the derived values of
``READ_DATA_RUN`` (src/fv_read_run.F90).

Lines are ``KEY = value``; ``!`` starts a comment; logicals are ``T``/``F``;
keys are case-sensitive as in the reference.
"""
from __future__ import annotations

import math

HARMONICS = ["M2", "S2", "N2", "K2", "K1", "O1", "P1", "Q1", "Mf", "Mm", "Ssa", "M4"]


def parse(text: str) -> dict:
    out = {}
    for line in text.splitlines():
        line = line.split("!", 1)[0]
        if "=" not in line:
            continue
        k, v = line.split("=", 1)
        k, v = k.strip(), v.strip()
        if k and k not in out:
            out[k] = v
    return out


class RunConfig:
    """Typed view of fv_run.dat plus the values READ_DATA_RUN derives."""

    def __init__(self, text: str):
        self.text = text
        self.raw = parse(text)

    # scalar accessors --------------------------------------------------
    def f(self, key, default=None):
        v = self.raw.get(key)
        if v is None:
            if default is None:
                raise KeyError(key)
            return float(default)
        return float(v.split()[0])

    def i(self, key, default=None):
        v = self.raw.get(key)
        if v is None:
            if default is None:
                raise KeyError(key)
            return int(default)
        return int(v.split()[0])

    def l(self, key, default=False):
        v = self.raw.get(key)
        if v is None:
            return bool(default)
        v = v.strip().upper()
        return v.startswith("T") or v.startswith(".T")

    def s(self, key, default=""):
        v = self.raw.get(key)
        return default if v is None else v.split()[0]

    # derived values ----------------------------------------------------
    @property
    def title(self):
        return self.s("TITLE")

    @property
    def is_LE(self):
        return self.title == "LE"

    @property
    def nsigma(self):
        return self.i("nsigma", 10)

    @property
    def Mt(self):
        """int(dt/dt_2D) (src/fv_read_run.F90:1060); an explicit ``Mt`` key overrides it."""
        if "Mt" in self.raw:
            return self.i("Mt")
        return int(self.f("dt") / self.f("dt_2D"))

    @property
    def type_task(self):
        t = 0
        if self.l("BAROTROPIC"):
            t = 1
        if self.l("BAROTROPIC_3D"):
            t = 2
        if self.l("SALINITY_ON") or self.l("TEMP_ON"):
            t = 3
        return t

    @property
    def harmonics(self):
        on = [0] * 12
        if self.l("TF_presence") or self.l("T_potential"):
            for tok in self.raw.get("Harmonics_tf", "").split():
                if tok in HARMONICS:
                    on[HARMONICS.index(tok)] = 1
        return on

    @property
    def n_harm(self):
        return sum(self.harmonics)

    @property
    def clim_relax(self):
        return 1.0 / (self.f("clim_relax", 15.0) * 3600.0 * 24.0)

    @property
    def relax2clim_ac(self):
        return 1.0 / (self.f("relax2clim_ac", 60.0) * 3600.0 * 24.0)

    @property
    def restart(self):
        return self.s("RESTART", "cold_start") == "hot_start"

    def unsupported(self):
        """Options outside the hot path this build implements (empty if none)."""
        bad = [k for k in ("mixed_BP_grad_OB", "T_potential",
                           "comp_sediment") if self.l(k)]
        if self.i("Vert_turb", 3) != 3:
            bad.append("Vert_turb")
        if self.l("key_atm") and self.i("type_swr_body", 2) == 1 and not self.l("jer_const", True):
            bad.append("jer_const")   # per-node Jerlov coefficients
        if self.i("type_swr_body", 2) not in (1, 2):   # src/fv_read_run.F90:799-802
            bad.append("type_swr_body")
        if self.type_task != 3:
            bad.append("type_task")
        return bad
