"""Statistics comparison shared by the GPU parity tests.

Per component m (n = occupancy, F = first order, S = second order):
  |dn_m|   <= rtol * n_m                       + g_hi * cnt_m     + eps_t * T_m
  |dF_md|  <= rtol * (sqrt(n_m S_md) + n_m r_d) + g_hi * sum|x_d|  + eps_t * T_m[|x_d|]
  |dS_md|  <= rtol * (S_md + n_m v_d)           + g_hi * sum x_d^2 + eps_t * T_m[x_d^2]
with v_d the frames' mean of x_d^2 and r_d = sqrt(v_d).  Scale: the
statistics are accumulated on frames centred by an offset o_d (the model's
weighted mean, or the frames' own mean in the mixed forms) and un-centred in
FP64, S = S' + 2 o F' + o^2 n, so their rounding scales with n (|x| + |o|)^2
rather than with S itself: an entry whose frames sit near x_d = 0 while o_d
is not small has an absolute error of u n o^2, far above u S (o_d^2 <= v_d).
Cut allowance (g_hi terms): cnt_m and the sums run over exactly the frames
whose FP64 posterior for component m lies within a relative band of the
1e-4 cut (oracle near_cut), g_hi = 1e-4 (1 + band).  A frame in the band may
land on either side of the cut in a finite-precision implementation, which
moves the statistics of m by at most its posterior times (1, |x|, x^2).
Bands: FP32 and fp16x3 1e-2, bf16x3 1e-1 (per-frame log-likelihood
tolerance 2e-4 relative, ~0.03 absolute at |LL| ~ 130).
Posterior allowance (eps_t terms): T_m = sum over frames of g(1-g) for the
posteriors g of m that pass the cut, and the same sums weighted by |x_d| and
x_d^2 (oracle tail_mass), with eps_t = 2e-4.  With every component
log-likelihood of a frame off by at most E (natural log), a posterior
g_m = exp(l_m - LL) moves by at most 2 E g_m (1 - g_m); eps_t = 2E for
E = 1e-4, a tenth of what the per-frame log-likelihood tolerance
(1e-5 relative at |LL| ~ 100) already allows.  A frame owned by one
component (g ~ 1) gets nothing; shared and tail posteriors get what a
1e-4 shift of the log-likelihoods does to them.  FP32 needs this: the
expanded contraction [1, x', x'^2] . [c, mu' ivar, -ivar/2] rounds FP32
partial sums of several hundred log2 units that cancel down to l2 (about
1e-5 per product), which reaches shared and tail posteriors undamped.
Without second-order statistics, F is scaled by n_m max|x| instead.
"""
import numpy as np
import pytest

import _oracle as orc

BAND = {"fp32": 1e-2, "fp16x3": 1e-2, "bf16x3": 1e-1}
EPS_TAIL = 2e-4


def near_cut(x, w, mu, var, band):
    return orc.near_cut(x, w, mu, var, band)


def check_stats(got, ref, x, w, mu, var, rtol=1e-5, tot_rtol=1e-6, band=1e-2, near=None,
                tail=None, what=""):
    """got/ref = (n, F, S or None, total LL or None).  Returns the largest
    fraction of any allowance used (for reports)."""
    n, F, S, tot = got[:4]
    rn, rF, rS, rtot = ref[:4]
    cnt, gx, gx2 = near if near is not None else near_cut(x, w, mu, var, band)
    tn, tx, tx2 = tail if tail is not None else orc.tail_mass(x, w, mu, var)
    ghi = orc.MIN_GAMMA * (1.0 + band)
    xd = np.asarray(x, np.float64)
    v = (xd * xd).mean(0) if xd.shape[0] else np.zeros(xd.shape[1])
    r = np.sqrt(v)
    if rS is not None:
        scaleF = np.sqrt(np.maximum(rn[:, None] * rS, 0.0)) + rn[:, None] * r[None]
    else:
        scaleF = rn[:, None] * (float(np.abs(xd).max()) if xd.size else 0.0)
    checks = [("n", n, rn, rtol * rn + ghi * cnt + EPS_TAIL * tn),
              ("F", F, rF, rtol * scaleF + ghi * gx + EPS_TAIL * tx)]
    if S is not None:
        checks.append(("S", S, rS, rtol * (rS + rn[:, None] * v[None]) + ghi * gx2
                       + EPS_TAIL * tx2))
    used = 0.0
    for name, g, rf, lim in checks:
        d = np.abs(np.asarray(g) - rf)
        bad = d > lim * (1 + 1e-12) + 1e-300
        if np.any(bad):
            idx = np.argwhere(bad)[:5]
            raise AssertionError(
                "%s%s out of tolerance at %s: got %s ref %s lim %s (near-cut %s, tail mass %s)" % (
                    what, name, idx.tolist(), np.asarray(g)[tuple(idx.T)], rf[tuple(idx.T)],
                    lim[tuple(idx.T)], cnt[idx[:, 0]], tn[idx[:, 0]]))
        pos = lim > 0
        if np.any(pos):
            used = max(used, float(np.max(d[pos] / lim[pos])))
    if tot is not None:
        assert tot == pytest.approx(rtot, rel=tot_rtol, abs=1e-6), (what, tot, rtot)
    return used
