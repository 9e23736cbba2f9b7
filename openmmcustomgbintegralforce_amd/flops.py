"""Algorithmic FLOP weights per interaction (fixed; never tuned to the kernels).

Convention: add/sub/mul = 1, FMA = 2, sqrt/rsqrt/rcp/div/exp/log = 1 each, counted from
the reference formulas (GBSWIntegral.cpp:176-182, 295-305; GBMVIntegralTypeI.cpp:164-171,
205-214; GBMVIntegralTypeII.cpp:174-189, 233-273; CharmmReferenceGBSW.cpp:248-282).
"""
C_VOL = {"gbsw": 16, "gbmv1": 12, "gbmv2": 23}   # per point/atom pair of the volume pass
C_FRC = {"gbsw": 32, "gbmv1": 28, "gbmv2": 60}   # per point/atom pair of the chain-rule pass
C_QP = 16      # per quadrature point (setup, cell index, accumulation)
C_PAIR = 40    # per unordered atom pair (STILL energy, dE/dr, dE/dB_i, dE/dB_j, force)
C_ATOM = 30    # per atom (Born map, self term, dE/dB -> dE/dI)


def volume_flops(kind: str, n_vol_pairs: float, n_points: float) -> float:
    return C_VOL[kind] * n_vol_pairs + C_QP * n_points


def force_flops(kind: str, n_frc_pairs: float) -> float:
    return C_FRC[kind] * n_frc_pairs


def pair_flops(n_atoms: int) -> float:
    return C_PAIR * n_atoms * (n_atoms - 1) / 2.0


def bin_bytes(n_binned: float, n_quad_points: float, n_atoms: int, slot_bytes: float = 16.0) -> float:
    """Algorithmic bytes of the binning pass: what a binned point occupies in its cell's slots
    (slot_bytes: 4 for a packed point, 16 for an FP32 record and packed point), a 4 B volume
    per point decided directly (outside the grid, empty cell, buried in its owner), 24 B of
    position per atom.  The brick records that binning by bricks writes and reads again are
    the method's own traffic, not algorithmic bytes."""
    return slot_bytes * n_binned + 4.0 * (n_quad_points - n_binned) + 24.0 * n_atoms
