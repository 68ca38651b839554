"""Algorithmic HBM bytes of the kernels whose roofline the benchmark reports.

Counting convention: compulsory traffic. Each distinct array a launch must read or
write counts once per launch (re-reads served by caches are not counted), except the
Gauss-Seidel sweeps, which count the per-contact gather/scatter of body velocities on
every sweep because every colour pass touches the bodies again.

Units: N bodies, E grid entries (covered cells summed over all boxes), G grid cells,
P emitted pairs, C contacts, S sweeps.
"""

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def pairs_bytes_per_body(dim):
    """k_pairs reads per body: sweep order (4), FP32 box (2D 16, 3D 24), exact FP64
    narrow data (2D q0, q1, r = 40; 3D q1, r = 32) and writes/reads one per-owner count or
    offset (4)."""
    box = 16 if dim == 2 else 24
    nar = 40 if dim == 2 else 32
    return 4 + box + nar + 4


def pairs_launch_bytes(dim, n, ncells, nentries, npairs, write_pass):
    """One k_pairs launch: per-body records, cell starts (4 per cell + 1), entries (4 each);
    the write pass adds the pair lists (8 per pair)."""
    b = n * pairs_bytes_per_body(dim) + 4 * (ncells + 1) + 4 * nentries
    if write_pass:
        b += 8 * npairs
    return b


def gs_sweep_bytes(dim, fp32):
    """Per contact per Gauss-Seidel sweep: the contact record (body ids, frame, Delassus
    diagonal, mu, forcing), the impulse read + write and both bodies' velocities (dof
    components) read + write."""
    w = 4 if fp32 else 8
    dof = 2 if dim == 2 else 6
    if dim == 2:
        rec = 8 + 2 * w + w + w + w  # ids, n, W, mu, forcing
    else:
        rec = 8 + 6 * w + 2 * w + 2 * w + w + w  # ids, n, s, arms, W_n, W_t, mu, forcing
    return rec + 2 * dim * w + 2 * 2 * dof * w


def gs_eval_bytes(dim, fp32):
    """Residual evaluation per contact: record and impulse read, both bodies read."""
    w = 4 if fp32 else 8
    dof = 2 if dim == 2 else 6
    return gs_sweep_bytes(dim, fp32) - dim * w - 2 * dof * w


def gs_launch_bytes(dim, ncontacts, sweeps, evals, fp32):
    return ncontacts * (sweeps * gs_sweep_bytes(dim, fp32) + evals * gs_eval_bytes(dim, fp32))


def broadphase_bytes(dim, n, npairs):
    """The broad phase of one step as a whole (predictor with keys, binning, cell ranges,
    neighbour sweep with the exact filter, ordered pair output), counted independently of
    how it is implemented: per body 52 + 48 + 16 + 48 + 8 = 172 B in 2D and
    72 + 48 + 12 + 68 + 8 = 208 B in 3D (FP64 state read once, FP32 boxes and radii,
    three 8-byte key/id radix passes, cell ranges, per-owner count and offset), plus 8 B
    per emitted pair."""
    return (172 if dim == 2 else 208) * n + 8 * npairs
