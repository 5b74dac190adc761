"""Drop-in GPU_WignerDirac2D_MassPotential_4x4 (GPU_WignerDirac2D_MassPotential_4x4.py:1620-3525).

The Wigner-Dirac 4x4 propagator with a position-dependent mass: the potential
factors use Mass(t, x) (the subclass's Mass_String) in place of mass/2
(:387-391, 422-427, 551-555, 585-591), the kinetic factors a fixed mass_half of
1e-4 (:1669), the antiparticle projectors the bare mass (:1668, 732-742).  Same
constructor, pre-set attributes and results as GPU_WignerDirac2D_4x4, plus
Mass_String.  The mass is tabulated on the host as the pair
F(x - theta/2) -/+ F(x + theta/2), F = Mass(x) c^2 dt, read by the fused p sweep.
"""
from __future__ import annotations

from .GPU_WignerDirac2D_4x4 import GPU_WignerDirac2D_4x4

KINETIC_MASS_HALF = 0.0001  # GPU_WignerDirac2D_MassPotential_4x4.py:1669


class GPU_WignerDirac2D_MassPotential_4x4(GPU_WignerDirac2D_4x4):
    """Wigner-Dirac propagator with the mass as a potential (Mass_String)."""

    def _make_plan(self, **variant):
        if not hasattr(self, "Mass_String"):
            raise AttributeError("set Mass_String before calling __init__")
        return super()._make_plan(mass_potential=self.Mass_String, kinetic_mass_half=KINETIC_MASS_HALF, **variant)
