"""The batched-ensemble workload: disks with uneven grids and parameters.

Draws (master seed 2026 by default, one SplitMix64 stream):
  grid        uniform over 128x128, 192x192, 256x256, 384x512, 512x512
  M_d/M*      U[0.3, 1.5]
  min Q       U[1.0, 2.0]
  softening   s ~ U[0.05, 0.2], eps = s^2 (the scale-free Kalnajs softening
              eps*R*R' of diskfft/README.md:51-55, i.e. a softening length s r)
  slope       U[-1.75, -1.25]
  noise seed  SplitMix64 output of (master seed + k)
Everything else is the power-law model's defaults (include/diskfft_model.h).
"""
import ctypes

GRIDS = [(128, 128), (192, 192), (256, 256), (384, 512), (512, 512)]


def _u(L, st):
    return (L.dfk_splitmix64(ctypes.byref(st)) >> 11) * (1.0 / 9007199254740992.0)


def draw(n, seed=2026, grids=GRIDS):
    import diskfft_amd as dfk
    L = dfk._model_lib()
    st = ctypes.c_ulonglong(seed)
    out = []
    for k in range(n):
        g = grids[min(int(_u(L, st) * len(grids)), len(grids) - 1)]
        q = 0.3 + 1.2 * _u(L, st)
        Qmin = 1.0 + 1.0 * _u(L, st)
        s = 0.05 + 0.15 * _u(L, st)
        slope = -1.75 + 0.5 * _u(L, st)
        ns = ctypes.c_ulonglong(seed + k)
        noise_seed = L.dfk_splitmix64(ctypes.byref(ns))
        out.append(dict(Nu=g[0], Nphi=g[1], mass_ratio=q, Qmin=Qmin, eps=s * s, slope=slope,
                        seed=noise_seed))
    return out


def model_for(dfk, d):
    return dfk.model_defaults(dfk.MODEL_POWERLAW, mass_ratio=d["mass_ratio"], Qmin=d["Qmin"],
                              slope=d["slope"], seed=d["seed"])


def build(dfk, disks, device=0):
    """Initial conditions of every disk and one handle holding all of them.
    Returns (Ensemble, [(Sig, vR, vphi, mD, params), ...])."""
    ics = []
    for d in disks:
        m = model_for(dfk, d)
        p, S, U, W, mD, _ = dfk.model_ic(m, d["Nu"], d["Nphi"], d["eps"], device=device)
        ics.append((S, U, W, mD, p))
    ens = dfk.Ensemble([ic[4] for ic in ics], device)
    for k, ic in enumerate(ics):
        ens.set_state(ic[0], ic[1], ic[2], d=k)
    return ens, ics
