"""Real-solid-harmonic AOs (test infrastructure only); filled in with the spherical basis path."""

from __future__ import annotations


def aos_sphe_value(aos, r_carts):
    raise NotImplementedError("spherical AOs are not restated yet")


def aos_sphe_vgl(aos, r_carts):
    raise NotImplementedError("spherical AOs are not restated yet")
