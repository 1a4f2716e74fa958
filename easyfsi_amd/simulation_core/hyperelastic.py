"""Neo-Hookean material constants (simulation_core/hyperelastic.py:26-180).

Host-side constants only: mu = E / (2 (1 + nu)), K = E / (3 (1 - 2 nu)),
lambda = K - 2 mu / 3; the Ecoflex 00-10 card calibrates mu from the 100 %
nominal modulus, mu = M100 / (2 - 2^-2).
"""

from __future__ import annotations

from dataclasses import dataclass
from math import sqrt

PSI_TO_PA = 6894.757293168361


def psi_to_pa(psi: float) -> float:
    return float(psi) * PSI_TO_PA


@dataclass(frozen=True)
class NeoHookeanMaterial:
    name: str
    density_kgm3: float
    shear_modulus_pa: float
    bulk_modulus_pa: float
    youngs_modulus_pa: float
    poissons_ratio: float

    @property
    def lame_lambda_pa(self) -> float:
        return self.bulk_modulus_pa - (2.0 / 3.0) * self.shear_modulus_pa

    @property
    def sound_speed_mps(self) -> float:
        return sqrt((self.bulk_modulus_pa + (4.0 / 3.0) * self.shear_modulus_pa) / self.density_kgm3)

    def stable_explicit_dt_s(self, spacing_m: float, cfl: float = 0.35) -> float:
        if spacing_m <= 0.0:
            raise ValueError("spacing_m must be positive")
        if cfl <= 0.0:
            raise ValueError("cfl must be positive")
        return float(cfl) * float(spacing_m) / self.sound_speed_mps

    @classmethod
    def from_youngs_modulus(cls, *, name: str, density_kgm3: float, youngs_modulus_pa: float,
                            poissons_ratio: float) -> "NeoHookeanMaterial":
        e = float(youngs_modulus_pa)
        nu = float(poissons_ratio)
        if e <= 0.0:
            raise ValueError("youngs_modulus_pa must be positive")
        if not 0.0 <= nu < 0.5:
            raise ValueError("poissons_ratio must be in [0, 0.5)")
        return cls(name=name, density_kgm3=float(density_kgm3), shear_modulus_pa=e / (2.0 * (1.0 + nu)),
                   bulk_modulus_pa=e / (3.0 * (1.0 - 2.0 * nu)), youngs_modulus_pa=e, poissons_ratio=nu)

    @classmethod
    def from_100_percent_modulus(cls, *, name: str, density_kgm3: float, modulus_100_pa: float,
                                 poissons_ratio: float = 0.49) -> "NeoHookeanMaterial":
        m100 = float(modulus_100_pa)
        if m100 <= 0.0:
            raise ValueError("modulus_100_pa must be positive")
        mu = m100 / (2.0 - 2.0 ** -2)
        return cls.from_youngs_modulus(name=name, density_kgm3=density_kgm3,
                                       youngs_modulus_pa=2.0 * mu * (1.0 + float(poissons_ratio)),
                                       poissons_ratio=poissons_ratio)


def ecoflex_0010_material(poissons_ratio: float = 0.49) -> NeoHookeanMaterial:
    """Ecoflex 00-10 starter card (density 1040, 100 % modulus 8 psi)."""
    return NeoHookeanMaterial.from_100_percent_modulus(name="Ecoflex 00-10", density_kgm3=1040.0,
                                                       modulus_100_pa=psi_to_pa(8.0), poissons_ratio=poissons_ratio)
