"""Writes vadv_inf_top.npz: a small vertical-advection case whose top level holds infinite u_stage values (and
infinite winds at the staggered level nk, which the reference never reads), with the expected utens_stage.

The expected values come from a pure-Python transcription of the reference functors u_forward_function /
u_backward_function (tests/regression/vertical_advection_dycore.cpp:32-122, BET_M = BET_P = 0.5 from
vertical_advection_defs.hpp:14-16), evaluated in Python floats (IEEE double, one rounding per operation, in the
source's order).  It is independent of the C oracle, which tests/test_oracle_vadv.py checks against this fixture.

    python tests/golden/make_vadv_inf_top.py      (rewrites the fixture next to this script)
"""
import os

import numpy as np

NI, NJ, NK, H = 7, 3, 6, 3
DTR = 0.15
BET_M = BET_P = 0.5


def inputs():
    rng = np.random.default_rng(20261203)
    shp = (NK, NJ + 2 * H, NI + 2 * H)
    uts, us, up, ut = (rng.uniform(-1, 1, shp) for _ in range(4))
    wc = rng.uniform(-0.5, 0.5, (NK + 1, NJ + 2 * H, NI + 2 * H))
    top = NK - 1
    # infinite u_stage at the top level in some columns (both signs), a NaN in one
    us[top, H + 0, H + 1] = np.inf
    us[top, H + 1, H + 3] = -np.inf
    us[top, H + 2, H + 5] = np.inf
    us[top, H + 2, H + 0] = np.nan
    # infinite winds on the staggered level nk, which no functor reads
    wc[NK, H + 0, H:H + NI + 1] = np.inf
    wc[NK, H + 1, H + 2] = -np.inf
    # one column with infinite u_stage one level below the top as well
    us[top - 1, H + 1, H + 6] = -np.inf
    return uts, us, wc, up, ut


def expected(uts, us, wc, up, ut):
    out = uts.copy()
    for j in range(H, H + NJ):
        for i in range(H, H + NI):
            U = lambda k: float(us[k, j, i])
            W = lambda di, k: float(wc[k, j, i + di])
            ccol = [0.0] * NK
            dcol = [0.0] * NK
            for k in range(NK):
                compute_d = DTR * float(up[k, j, i]) + float(ut[k, j, i]) + float(uts[k, j, i])
                if k == 0:  # apply(first_level)
                    gcv = 0.25 * (W(1, k + 1) + W(0, k + 1))
                    cs = gcv * BET_M
                    c = gcv * BET_P
                    b = DTR - c
                    correction = -cs * (U(k + 1) - U(k))
                    d = compute_d + correction
                    divided = 1.0 / b
                    ccol[k] = c * divided
                    dcol[k] = d * divided
                elif k == NK - 1:  # apply(last_level)
                    gav = -0.25 * (W(1, k) + W(0, k))
                    as_ = gav * BET_M
                    a = gav * BET_P
                    b = DTR - a
                    correction = -as_ * (U(k - 1) - U(k))
                    d = compute_d + correction
                    divided = 1.0 / (b - ccol[k - 1] * a)
                    dcol[k] = (d - dcol[k - 1] * a) * divided
                else:  # apply(modify<1, -1>)
                    gav = -0.25 * (W(1, k) + W(0, k))
                    gcv = 0.25 * (W(1, k + 1) + W(0, k + 1))
                    as_ = gav * BET_M
                    cs = gcv * BET_M
                    a = gav * BET_P
                    c = gcv * BET_P
                    b = DTR - a - c
                    correction = -as_ * (U(k - 1) - U(k)) - cs * (U(k + 1) - U(k))
                    d = compute_d + correction
                    divided = 1.0 / (b - ccol[k - 1] * a)
                    ccol[k] = c * divided
                    dcol[k] = (d - dcol[k - 1] * a) * divided
            data_col = 0.0
            for k in range(NK - 1, -1, -1):
                if k == NK - 1:  # u_backward_function::apply(last_level)
                    out[k, j, i] = DTR * (dcol[k] - float(up[k, j, i]))
                    data_col = dcol[k]
                else:
                    data = dcol[k] - ccol[k] * data_col
                    out[k, j, i] = DTR * (data - float(up[k, j, i]))
                    data_col = data
    return out


def main():
    with np.errstate(invalid="ignore", over="ignore"):
        uts, us, wc, up, ut = inputs()
        ref = expected(uts, us, wc, up, ut)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vadv_inf_top.npz")
    np.savez_compressed(path, utens_stage=uts, u_stage=us, wcon=wc, u_pos=up, utens=ut, expected=ref,
                        dims=np.array([NI, NJ, NK, H]), dtr=np.array(DTR))
    print(f"wrote {path}: {np.isinf(ref).sum()} inf, {np.isnan(ref).sum()} nan of {ref[:, H:H + NJ, H:H + NI].size}")


if __name__ == "__main__":
    main()
