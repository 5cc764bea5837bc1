"""Dispersive material models: closed-form permittivity and ADE coefficients.

Formulas are the reference's (src/prismo/materials/dispersion.py):
Lorentz eps(w) = eps_inf + sum de w0^2 / (w0^2 - w^2 - i w g) (177-187) and its
bilinear ADE coefficients (189-231); Drude eps(w) = eps_inf - wp^2/(w^2 + i w g)
(260-265) and its exponential-integrator coefficients (267-286); Debye
(315-336); Sellmeier (361-386).  ``discrete_susceptibility`` is the exact
z-domain response of those recursions once they are coupled into the E
update the way the stencil kernels couple them.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass

import numpy as np


@dataclass
class LorentzPole:
    omega_0: float
    delta_epsilon: float
    gamma: float


@dataclass
class DrudePole:
    omega_p: float
    gamma: float


@dataclass
class DebyePole:
    epsilon_s: float
    tau: float


class DispersiveMaterial(ABC):
    def __init__(self, epsilon_inf: float = 1.0, name: str = ""):
        self.epsilon_inf = epsilon_inf
        self.name = name or self.__class__.__name__

    @abstractmethod
    def permittivity(self, omega): ...

    def refractive_index(self, omega):
        return np.sqrt(self.permittivity(omega))

    @abstractmethod
    def get_ade_coefficients(self, dt: float) -> dict: ...


class LorentzMaterial(DispersiveMaterial):
    def __init__(self, epsilon_inf: float, poles, name: str = ""):
        super().__init__(epsilon_inf, name)
        self.poles = list(poles)

    def permittivity(self, omega):
        eps = self.epsilon_inf * np.ones_like(omega, dtype=complex)
        for p in self.poles:
            eps += p.delta_epsilon * p.omega_0**2 / (p.omega_0**2 - omega**2 - 1j * omega * p.gamma)
        return eps

    def get_ade_coefficients(self, dt: float) -> dict:
        poles = []
        for p in self.poles:
            w0, g, de = p.omega_0, p.gamma, p.delta_epsilon
            denom = 4.0 + 2 * g * dt + w0**2 * dt**2
            c0 = 2 * de * w0**2 * dt**2 / denom
            poles.append({
                "C0": c0,
                "C1": c0,
                "C2": (8.0 - 2 * w0**2 * dt**2) / denom,
                "C3": -(4.0 - 2 * g * dt + w0**2 * dt**2) / denom,
                "omega_0": w0, "gamma": g, "delta_epsilon": de,
            })
        return {"poles": poles, "epsilon_inf": self.epsilon_inf}


class DrudeMaterial(DispersiveMaterial):
    def __init__(self, epsilon_inf: float, omega_p: float, gamma: float, name: str = ""):
        super().__init__(epsilon_inf, name)
        self.omega_p = omega_p
        self.gamma = gamma

    def permittivity(self, omega):
        return self.epsilon_inf - self.omega_p**2 / (omega**2 + 1j * omega * self.gamma)

    def get_ade_coefficients(self, dt: float) -> dict:
        e = np.exp(-self.gamma * dt)
        return {
            "C0": self.omega_p**2 / self.gamma * (1.0 - e),
            "C1": e,
            "omega_p": self.omega_p, "gamma": self.gamma, "epsilon_inf": self.epsilon_inf,
        }


class DebyeMaterial(DispersiveMaterial):
    def __init__(self, epsilon_inf: float, epsilon_s: float, tau: float, name: str = ""):
        super().__init__(epsilon_inf, name)
        self.epsilon_s = epsilon_s
        self.tau = tau

    def permittivity(self, omega):
        return self.epsilon_inf + (self.epsilon_s - self.epsilon_inf) / (1.0 + 1j * omega * self.tau)

    def get_ade_coefficients(self, dt: float) -> dict:
        e = np.exp(-dt / self.tau)
        return {
            "C0": (self.epsilon_s - self.epsilon_inf) * (1.0 - e),
            "C1": e,
            "tau": self.tau, "epsilon_s": self.epsilon_s, "epsilon_inf": self.epsilon_inf,
        }


class SellmeierMaterial(DispersiveMaterial):
    def __init__(self, B_coeffs, C_coeffs, name: str = ""):
        super().__init__(epsilon_inf=1.0, name=name)
        self.B = np.array(B_coeffs)
        self.C = np.array(C_coeffs)

    def permittivity(self, omega):
        lam_um = 2 * np.pi * 299792458.0 / omega * 1e6
        l2 = lam_um**2
        n2 = 1.0
        for b, c in zip(self.B, self.C):
            n2 += b * l2 / (l2 - c)
        return n2

    def get_ade_coefficients(self, dt: float) -> dict:
        raise NotImplementedError(
            "Sellmeier model should be fitted to Lorentz poles for time-domain simulation"
        )


def discrete_susceptibility(material, dt: float, omega):
    """Exact response eps_num(w) = eps_inf + chi(z) of the coupled
    recursions the E kernels run, z^-1 = exp(i w dt) (time convention
    exp(-i w t), as the closed forms above).

    The kernels feed the coefficients above with the drive averaged over the
    half steps on either side of each update:
    Lorentz: P+ = C2 P + C3 P- + [(C0 E+ + C1 E) + (C0 E + C1 E-)] / 2
             -> chi = (C0 + C1 z^-1)(1 + z^-1) / 2 / (1 - C2 z^-1 - C3 z^-2)
    Drude:   J+ = C1 J + C0 (E+ + E) / 2,  P+ - P = dt (J+ + J) / 2
             -> chi = dt/2 (1 + z^-1)/(1 - z^-1) * C0 (1 + z^-1) / 2 / (1 - C1 z^-1)
    Both keep chi(w -> 0) of the closed form and differ from it by
    O((w dt)^2); the uncentred use (drive C0 E+ + C1 E alone) would carry an
    extra phase exp(-i w dt / 2), i.e. a spurious gain.
    """
    zi = np.exp(1j * np.asarray(omega) * dt)  # z^-1 for exp(-i w t)
    if isinstance(material, LorentzMaterial):
        chi = 0
        for p in material.get_ade_coefficients(dt)["poles"]:
            chi = chi + 0.5 * (p["C0"] + p["C1"] * zi) * (1 + zi) / (
                1 - p["C2"] * zi - p["C3"] * zi**2)
        return material.epsilon_inf + chi
    if isinstance(material, DrudeMaterial):
        co = material.get_ade_coefficients(dt)
        chi = 0.25 * dt * (1 + zi) ** 2 * co["C0"] / ((1 - co["C1"] * zi) * (1 - zi))
        return material.epsilon_inf + chi
    raise TypeError(f"no discrete model for {type(material).__name__}")
