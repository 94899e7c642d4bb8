"""Algorithmic work per unit, frozen here and used by bench.py.

Counting convention: add, sub, mul, div, sqrt, log, pow = 1 flop each, a
fused multiply-add = 2; parameter-only subexpressions (G m, b^2, ...) are
hoisted and not counted.  Counts follow the reference formulas:

  Hernquist grad   builtin_potentials.cpp:221-233        16
  Plummer grad     :302-314                              15
  Miyamoto-Nagai   :1298-1312                            20
  spherical NFW    :835-850                              22
  composite (nucleus, bulge, disk, halo)                 74
  leapfrog update  leapfrog.pyx:35-51                    18  -> 92 per orbit-step
  DOP853 accepted  dop853.cpp:369-459: 12 evaluations (12 x 74) + stage
                   combinations 666 + b-combination 102 + error norm 174 +
                   control 10                           ~1,840
  DOP853 rejected  11 evaluations, no FSAL              ~1,766
  SCF Phi + grad   (nmax 10, lmax 6; 308 active terms): 8 T + 20 L + 12 R
                   + 8 L + 50 with T = 308, L = 28, R = 77 ~4,222
  MLP forward      2 (4*256 + 3*256^2 + 256*4)         397,312 (BF16 MFMA)

These are ALGORITHMIC counts: transcendental functions (sqrt, divide, log)
expand to many FP64 instructions on CDNA, so the instruction rate rocprof
reports is higher than these flops imply.
"""

FLOP_COMPOSITE_GRAD = 74
FLOP_LEAPFROG_STEP = 92
FLOP_DOP853_ACCEPTED = 1840
FLOP_DOP853_REJECTED = 1766
FLOP_DOP853_DENSE_EXTRA = 1140
FLOP_SCF_GRAD = 4222
FLOP_SCF_LEAPFROG_STEP = 4240
FLOP_MLP_FORWARD = 2 * (4 * 256 + 3 * 256 * 256 + 256 * 4)
# the orbit force of the surrogate: Phi_net's forward (o0 only) and the
# reverse pass for d o0 / d features (three hidden layers and the first
# layer's 4 x 256 reduction); the matrix cores execute the hidden products
# three times (hi*hi + hi*lo + lo*hi of the BF16 split), not counted here
FLOP_MLP_FORCE = 2 * (4 * 256 + 3 * 256 * 256 + 256) + 2 * (3 * 256 * 256 + 4 * 256)
# mock-stream particle-step: composite gradient 74, shift to the progenitor
# 3, Plummer gradient 15, leapfrog update 18 (mockstream.pyx:584-599)
FLOP_STREAM_PARTICLE_STEP = 110
# bytes a saved leapfrog step writes per orbit (six doubles)
BYTES_SAVED_STATE = 48
BYTES_LEAPFROG_SAVE = 48  # one (x, v) state written per saved orbit-step

# MI355X published peaks
PEAK_FP64_TFLOPS = 78.6      # FP64 vector (and FP64 matrix) dense
PEAK_BF16_TFLOPS = 2500.0    # BF16 dense MFMA
PEAK_HBM_GBS = 8000.0


def dop853_flops(naccpt, nrejct):
    return naccpt * FLOP_DOP853_ACCEPTED + nrejct * FLOP_DOP853_REJECTED


def copy_calibration(n_doubles=1 << 25, reps=5):
    """Run a plain elementwise kernel of known traffic (torch.neg on
    n_doubles FP64 values: 8 n bytes read, 8 n bytes written per launch) so
    that rocprofv3's FETCH_SIZE / WRITE_SIZE counters can be scaled to
    bytes: factor = known bytes / counter bytes of this kernel."""
    import torch
    src = torch.ones(n_doubles, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(f"copy calibration: {reps} launches of torch.neg over {n_doubles} doubles "
          f"({8 * n_doubles} bytes read, {8 * n_doubles} bytes written each)", flush=True)


if __name__ == "__main__":
    import sys
    if len(sys.argv) > 1 and sys.argv[1] == "calibrate":
        copy_calibration()
    else:
        sys.exit("usage: python -m gala_amd.roofline calibrate")
