"""CPU reference of the device Metropolis loop (test infrastructure).

Philox4x32-10 restated in numpy from the published algorithm (Salmon et al.,
"Parallel random numbers: as easy as 1, 2, 3", SC'11), Box-Muller normals,
and the chain step of cpab/cpaNd/inference/Metropolis.py:41-85 with the
oracle's calc_T_inv, remap and SSD likelihood.
"""
import numpy as np

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF
SLOT_ACCEPT = 0x80000000


def philox4x32_10(ctr, key):
    c0, c1, c2, c3 = [int(x) & MASK for x in ctr]
    k0, k1 = [int(x) & MASK for x in key]
    for _ in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> 32, p0 & MASK
        hi1, lo1 = p1 >> 32, p1 & MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        k0 = (k0 + W0) & MASK
        k1 = (k1 + W1) & MASK
    return c0, c1, c2, c3


def u53(a, b):
    return ((a >> 5) * 67108864.0 + (b >> 6)) * (1.0 / 9007199254740992.0)


def normals(step, chain, d, seed):
    z = np.zeros(d)
    for t in range((d + 1) // 2):
        r = philox4x32_10((step & MASK, step >> 32, chain, t), (seed & MASK, seed >> 32))
        u1 = 1.0 - u53(r[0], r[1])
        u2 = u53(r[2], r[3])
        rad = np.sqrt(-2.0 * np.log(u1))
        ang = 6.283185307179586 * u2
        z[2 * t] = rad * np.cos(ang)
        if 2 * t + 1 < d:
            z[2 * t + 1] = rad * np.sin(ang)
    return z


def accept_uniform(step, chain, seed):
    r = philox4x32_10((step & MASK, step >> 32, chain, SLOT_ACCEPT), (seed & MASK, seed >> 32))
    return u53(r[0], r[1])


def log_likelihood(orc, space, B, theta, src, dst, sigma, params, use_ave_ll=True):
    n = src.shape[0]
    Av = B @ theta
    sA, T = orc.prep(2, Av.reshape(space.nC, -1), -1, params["dt"])
    grid = space.tessellation.dense_grid_img()
    inv = orc.calc_T(space.tessellation.geom(), grid, T, sA, params["dt"], params["nTimeSteps"],
                     params["nStepsODEsolver"])
    warped = orc.remap2d(inv, src)
    ll = orc.ssd_ll(warped, dst, sigma)
    return ll / (n * n) if use_ave_ll else ll


def log_prior(J, theta):
    """lp = -0.5 theta^T J theta (cpab/cpaNd/model/_ScaleDependentLogPrior.py,
    mu = 0): (theta^T J)_j summed over i in index order, then the products
    theta_j (theta^T J)_j summed in index order, in IEEE double."""
    d = len(theta)
    q = 0.0
    for j in range(d):
        v = 0.0
        for i in range(d):
            v += float(theta[i]) * float(J[i, j])
        q += v * float(theta[j])
    return -0.5 * q


def run_chain(orc, space, B, L, theta0, src, dst, sigma, params, nsteps, seed, chain=0,
              use_ave_ll=True, inverted=False, J=None, wlp=0.0):
    """Returns per-step (accepted, theta after the step, ll after the step, proposal ll).
    With a prior precision J the log ratio gains wlp (lp' - lp)
    (cpab/cpaNd/inference/Metropolis.py:64-70)."""
    theta = np.array(theta0, dtype=np.float64)
    ll = log_likelihood(orc, space, B, theta, src, dst, sigma, params, use_ave_ll)
    lp = log_prior(J, theta) if J is not None else 0.0
    out = []
    for step in range(nsteps):
        z = normals(step, chain, len(theta), seed)
        prop = theta + np.array([sum(L[i, j] * z[j] for j in range(i + 1)) for i in range(len(theta))])
        llp = log_likelihood(orc, space, B, prop, src, dst, sigma, params, use_ave_ll)
        lr = llp - ll
        if J is not None:
            lpp = log_prior(J, prop)
            lr += wlp * (lpp - lp)
        if lr > 0:
            acc = True
        else:
            lu = np.log(accept_uniform(step, chain, seed))
            acc = (lu > lr) if inverted else (lu < lr)
        if acc:
            theta, ll = prop, llp
            if J is not None:
                lp = lpp
        out.append((acc, theta.copy(), ll, llp))
    return out
