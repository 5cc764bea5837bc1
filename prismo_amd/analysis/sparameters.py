"""S-parameters from port flux monitors (API of
src/prismo/analysis/sparameters.py:14-279).

``SParameterAnalyzer`` keeps the reference's interface and definitions:
S_ij = sqrt(P_i / P_inc,j) for transmission, S_ii = sqrt(P_back / P_inc)
(sparameters.py:47-86), an (n_freq, n_port, n_port) matrix, insertion and
return loss, reciprocity/unitarity checks, Touchstone export
(sparameters.py:233-279) and group delay.

``port_power`` supplies the power those definitions need: the time-averaged
Poynting flux through a port line, normal to it only,
  P = 1/2 Re sum_cells (E x H*) . n dA,
from a FieldMonitor's spectra.  (The reference's FieldMonitor.get_power_flow
adds every E x H product whatever the port normal, field.py:268-287, which
is not a port power.)
"""

from __future__ import annotations

from pathlib import Path
from typing import Optional

import numpy as np

# (E, H, sign) pairs of (E x H) . n for each port normal
_NORMAL_PAIRS = {
    "x": (("Ey", "Hz", 1.0), ("Ez", "Hy", -1.0)),
    "y": (("Ez", "Hx", 1.0), ("Ex", "Hz", -1.0)),
    "z": (("Ex", "Hy", 1.0), ("Ey", "Hx", -1.0)),
}


def port_power(monitor, normal: str = "x", direction: int = 1, frequencies=None) -> np.ndarray:
    """Flux of 1/2 Re(E x H*) through the monitor's cells along +/- ``normal``,
    one value per frequency (W per metre in 2-D).  Only the pairs the monitor
    records enter; E and H sample points are taken as coincident cell by cell."""
    freqs = monitor.frequencies if frequencies is None else frequencies
    g = monitor._grid
    dx, dy, dz = g.spacing
    area = {"x": dy * (1.0 if g.is_2d else dz), "y": dx * (1.0 if g.is_2d else dz),
            "z": dx * dy}[normal]
    out = np.zeros(len(freqs))
    stacked = frequencies is None and hasattr(monitor, "frequency_stack")

    def spectra(c):
        if stacked:
            a = monitor.frequency_stack(c)
            return a.reshape(a.shape[0], -1)
        return np.stack([np.asarray(monitor.get_frequency_data(c, f)).ravel() for f in freqs])

    for e, h, sgn in _NORMAL_PAIRS[normal]:
        if e in monitor.components and h in monitor.components:
            E = spectra(e)
            H = spectra(h)
            n = min(E.shape[1], H.shape[1])
            out += sgn * 0.5 * np.real(np.sum(E[:, :n] * np.conj(H[:, :n]), axis=1)) * area
    return direction * out


class SParameterAnalyzer:
    def __init__(self, num_ports: int, frequencies, reference_impedance: float = 50.0):
        self.num_ports = num_ports
        self.frequencies = np.asarray(frequencies)
        self.z0 = reference_impedance
        self.s_matrix = np.zeros((len(self.frequencies), num_ports, num_ports), dtype=complex)
        self.port_powers: dict = {}

    def add_port_data(self, port_index: int, excitation_port: int, power_forward,
                      power_backward) -> None:
        """sparameters.py:47-86: the first forward power recorded for an
        excitation is its incident power."""
        if excitation_port not in self.port_powers:
            self.port_powers[excitation_port] = np.asarray(power_forward, dtype=float)
        inc = self.port_powers[excitation_port]
        if port_index == excitation_port:
            self.s_matrix[:, port_index, excitation_port] = (np.asarray(power_backward) / inc) ** 0.5
        else:
            self.s_matrix[:, port_index, excitation_port] = (np.asarray(power_forward) / inc) ** 0.5

    def add_mode_data(self, port_index: int, excitation_port: int, mode_coefficients: dict) -> None:
        """The reference's definitions (sparameters.py:88-122): reflection
        S_ii = backward / forward, transmission S_ij = forward / backward of
        the coefficients at port i.  Added: with an ``'incident'`` key (the
        exciting port's forward coefficient), transmission is forward /
        incident instead, the usual mode-expansion S_ij; without it the
        reference's quotient is kept."""
        fwd = mode_coefficients.get("forward", np.zeros_like(self.frequencies))
        bwd = mode_coefficients.get("backward", np.zeros_like(self.frequencies))
        if port_index == excitation_port:
            self.s_matrix[:, port_index, excitation_port] = bwd / fwd
        elif "incident" in mode_coefficients:
            self.s_matrix[:, port_index, excitation_port] = fwd / mode_coefficients["incident"]
        else:
            self.s_matrix[:, port_index, excitation_port] = fwd / bwd

    def get_s_parameter(self, i: int, j: int) -> np.ndarray:
        return self.s_matrix[:, i, j]

    def get_s_matrix(self, frequency_index: int) -> np.ndarray:
        return self.s_matrix[frequency_index, :, :]

    def get_insertion_loss_db(self, i: int, j: int) -> np.ndarray:
        return -20 * np.log10(np.abs(self.s_matrix[:, i, j]))

    def get_return_loss_db(self, port: int) -> np.ndarray:
        return -20 * np.log10(np.abs(self.s_matrix[:, port, port]))

    def check_reciprocity(self) -> float:
        err = 0.0
        for i in range(self.num_ports):
            for j in range(i + 1, self.num_ports):
                err = max(err, float(np.max(np.abs(self.s_matrix[:, i, j] - self.s_matrix[:, j, i]))))
        return err

    def check_unitarity(self, frequency_index: int) -> float:
        S = self.s_matrix[frequency_index]
        return float(np.max(np.abs(S.conj().T @ S - np.eye(self.num_ports))))


def export_touchstone(filename, frequencies, s_matrix, z0: float = 50.0,
                      comments: Optional[list] = None) -> None:
    """The reference's Touchstone layout (sparameters.py:233-279): comment
    lines, '# Hz S RI R z0', then per frequency the S_ij real/imaginary
    pairs in row-major port order."""
    s_matrix = np.asarray(s_matrix)
    n = s_matrix.shape[1]
    with open(Path(filename), "w") as f:
        f.write("! Touchstone file generated by prismo_amd\n")
        for c in comments or []:
            f.write(f"! {c}\n")
        f.write(f"# Hz S RI R {z0}\n")
        for k, freq in enumerate(np.asarray(frequencies)):
            f.write(f"{freq:.10e}")
            for i in range(n):
                for j in range(n):
                    v = s_matrix[k, i, j]
                    f.write(f" {v.real:.10e} {v.imag:.10e}")
            f.write("\n")


def compute_group_delay(frequencies, s_parameter) -> np.ndarray:
    """-d(phase)/d(omega) of an S-parameter (unwrapped phase)."""
    phase = np.unwrap(np.angle(np.asarray(s_parameter)))
    return -np.gradient(phase, 2 * np.pi * np.asarray(frequencies))
