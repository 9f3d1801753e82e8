"""Writes tests/golden/*.npz: small rooms evaluated by an independent numpy implementation.

This evaluation is written separately from the C oracle: vectorised over whole slices with
numpy, classification by array shifts, and its own step loop. It follows the same reference
files (kernels3d.cu:485-529 for the update body, :83-191 for the step order, cudaMesh.cu:328-361
for the classification meaning). numpy elementwise arithmetic performs one IEEE rounding per
operation, like the oracle built with -ffp-contract=off, so the fixtures are exact.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent


def classify(lab: np.ndarray):
    """pos/mat bytes from labels (0 inside, m+1 solid of material m)."""
    Z, Y, X = lab.shape
    pad = np.ones((Z + 2, Y + 2, X + 2), np.int32)  # outside the grid: solid, material 0
    pad[1:-1, 1:-1, 1:-1] = lab
    inside = lab == 0
    # neighbours in the order -x, +x, -y, +y, -z, +z
    nbs = [pad[1:-1, 1:-1, :-2], pad[1:-1, 1:-1, 2:], pad[1:-1, :-2, 1:-1], pad[1:-1, 2:, 1:-1],
           pad[:-2, 1:-1, 1:-1], pad[2:, 1:-1, 1:-1]]
    K = sum((n == 0).astype(np.int32) for n in nbs)
    m = np.zeros(lab.shape, np.int32)
    found = np.zeros(lab.shape, bool)
    for n in nbs:
        take = (~found) & (n != 0)
        m[take] = n[take] - 1
        found |= take
    pos = np.where(inside, 0x80 | K, 0).astype(np.uint8)
    mat = np.where(inside & (K < 6), m, 0).astype(np.uint8)
    return pos, mat


def run(pos, mat, betas, octave, lam, src, kinds, sig, rec, n_steps, dt, literal=0):
    """Step loop; betas[m] is the admittance the update uses for material byte m."""
    Z, Y, X = pos.shape
    one = dt(1)
    half = 0.5
    if dt == np.float32:
        lamv, lam2 = np.float32(lam), np.float32(lam * lam)
    else:
        lamv, lam2 = lam, lam * lam
    K = (pos & 0x7F).astype(dt)
    sw = (pos >> 7).astype(dt)
    b = np.asarray(betas, dt)[mat]
    # beta_e = 0.5 * (beta * (6 - K) * lambda); in FP32 the 0.5 product is taken in double
    prod = (b * (dt(6) - K)) * lamv
    beta = (half * prod.astype(np.float64)).astype(dt) if dt == np.float32 else half * prod
    inv = one / (one + beta)
    a = sw * inv
    tk = dt(2) - K * lam2
    omb = one - beta
    P = np.zeros((Z, Y, X), dt)
    Pp = np.zeros((Z, Y, X), dt)
    resp = np.zeros((len(rec), n_steps), dt)
    for n in range(n_steps):
        for s, (x, y, z) in enumerate(src):
            v = dt(sig[s, n])
            if kinds[s] == 0 or (literal & 2):
                P[z, y, x] = v
            else:
                P[z, y, x] = P[z, y, x] + v
        c = (slice(1, Z - 1), slice(1, Y - 1), slice(1, X - 1))
        S = P[2:, 1:-1, 1:-1] + P[:-2, 1:-1, 1:-1]
        S = S + P[1:-1, 2:, 1:-1]
        S = S + P[1:-1, :-2, 1:-1]
        S = S + P[1:-1, 1:-1, 2:]
        S = S + P[1:-1, 1:-1, :-2]
        new = a[c] * ((tk[c] * P[c] + lam2 * S) - omb[c] * Pp[c])
        Pn = Pp
        Pn[1:-1, :, :] = 0
        Pn[c] = new
        P, Pp = Pn, P
        for r, (x, y, z) in enumerate(rec):
            resp[r, n] = P[z, y, x]
    return resp, P, Pp


def run_f16(pos, mat, betas, lam, src, kinds, sig, rec, n_steps, scale_exp, literal=0,
            diff=False):
    """FP32 arithmetic with binary16 field storage scaled by 2^scale_exp: the plain storage of
    (P, P_past), or (diff) the difference form (P, D = P - P_past): the past value is P - D, each
    update stores fl16(new) and fl16(new - P), and a source sample keeps the past value,
    D = fl16(P_new - (P_old - D_old))."""
    Z, Y, X = pos.shape
    f = np.float32
    r16 = lambda v: np.asarray(v, f).astype(np.float16).astype(f)
    lamv, lam2 = f(lam), f(lam * lam)
    K = (pos & 0x7F).astype(f)
    sw = (pos >> 7).astype(f)
    b = np.asarray(betas, f)[mat]
    beta = (0.5 * ((b * (f(6) - K)) * lamv).astype(np.float64)).astype(f)
    a = sw * (f(1) / (f(1) + beta))
    tk = f(2) - K * lam2
    omb = f(1) - beta
    scale, inv_scale = f(2.0 ** scale_exp), f(2.0 ** -scale_exp)
    P = np.zeros((Z, Y, X), f)
    B = np.zeros((Z, Y, X), f)  # P_past, or D in the difference form
    resp = np.zeros((len(rec), n_steps), f)
    c = (slice(1, Z - 1), slice(1, Y - 1), slice(1, X - 1))
    for n in range(n_steps):
        for s, (x, y, z) in enumerate(src):
            v = f(sig[s, n]) * scale
            hard = kinds[s] == 0 or (literal & 2)
            past = P[z, y, x] - B[z, y, x]
            P[z, y, x] = r16(v) if hard else r16(P[z, y, x] + v)
            if diff:
                B[z, y, x] = r16(P[z, y, x] - past)
        S = P[2:, 1:-1, 1:-1] + P[:-2, 1:-1, 1:-1]
        S = S + P[1:-1, 2:, 1:-1]
        S = S + P[1:-1, :-2, 1:-1]
        S = S + P[1:-1, 1:-1, 2:]
        S = S + P[1:-1, 1:-1, :-2]
        past = P[c] - B[c] if diff else B[c]
        new = a[c] * ((tk[c] * P[c] + lam2 * S) - omb[c] * past)
        Pn = np.zeros((Z, Y, X), f)  # slices 0 and Z-1 and the x/y edges stay 0
        Pn[c] = r16(new)
        if diff:
            Bn = np.zeros((Z, Y, X), f)
            Bn[c] = r16(new - P[c])
            P, B = Pn, Bn
        else:
            P, B = Pn, P
        for r, (x, y, z) in enumerate(rec):
            resp[r, n] = P[z, y, x] * inv_scale
    return resp, P, B


def run_sumdiff16(pos, mat, betas, lam, src, kinds, sig, rec, n_steps, scale_exp, literal=0):
    """The sum-difference form: binary16 S = P + P_past and D = P - P_past (scaled), P and P_past
    reconstructed as (S + D)/2 and (S - D)/2 in FP32; each update stores fl16(new + P) and
    fl16(new - P); a source sample keeps the past value; receivers read (S + D)/2."""
    Z, Y, X = pos.shape
    f = np.float32
    r16 = lambda v: np.asarray(v, f).astype(np.float16).astype(f)
    lamv, lam2 = f(lam), f(lam * lam)
    K = (pos & 0x7F).astype(f)
    sw = (pos >> 7).astype(f)
    b = np.asarray(betas, f)[mat]
    beta = (0.5 * ((b * (f(6) - K)) * lamv).astype(np.float64)).astype(f)
    a = sw * (f(1) / (f(1) + beta))
    tk = f(2) - K * lam2
    omb = f(1) - beta
    scale, inv_scale, h = f(2.0 ** scale_exp), f(2.0 ** -scale_exp), f(0.5)
    S = np.zeros((Z, Y, X), f)
    D = np.zeros((Z, Y, X), f)
    resp = np.zeros((len(rec), n_steps), f)
    c = (slice(1, Z - 1), slice(1, Y - 1), slice(1, X - 1))
    for n in range(n_steps):
        for s, (x, y, z) in enumerate(src):
            v = f(sig[s, n]) * scale
            p = (S[z, y, x] + D[z, y, x]) * h
            past = (S[z, y, x] - D[z, y, x]) * h
            pn = v if (kinds[s] == 0 or (literal & 2)) else p + v
            S[z, y, x], D[z, y, x] = r16(pn + past), r16(pn - past)
        P = (S + D) * h
        Pp = (S - D) * h
        T = P[2:, 1:-1, 1:-1] + P[:-2, 1:-1, 1:-1]
        T = T + P[1:-1, 2:, 1:-1]
        T = T + P[1:-1, :-2, 1:-1]
        T = T + P[1:-1, 1:-1, 2:]
        T = T + P[1:-1, 1:-1, :-2]
        new = a[c] * ((tk[c] * P[c] + lam2 * T) - omb[c] * Pp[c])
        Sn = np.zeros((Z, Y, X), f)
        Dn = np.zeros((Z, Y, X), f)
        Sn[0], Sn[-1], Dn[0], Dn[-1] = S[0], S[-1], D[0], D[-1]
        Sn[c] = r16(new + P[c])
        Dn[c] = r16(new - P[c])
        S, D = Sn, Dn
        for r, (x, y, z) in enumerate(rec):
            resp[r, n] = (S[z, y, x] + D[z, y, x]) * h * inv_scale
    return resp, S, D


def splitmix(seed):
    state = seed & (2**64 - 1)
    while True:
        state = (state + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        yield z ^ (z >> 31)


def case(seed: int, dims, n_mat: int, n_steps: int, literal: int, octave: int):
    g = splitmix(seed)
    u = lambda: (next(g) >> 11) * (1.0 / (1 << 53))
    ri = lambda lo, hi: lo + next(g) % (hi - lo + 1)
    X, Y, Z = dims
    lab = np.zeros((Z, Y, X), np.uint8)
    lab[:, :, 0] = lab[:, :, -1] = lab[:, 0, :] = lab[:, -1, :] = 1 + min(2, n_mat - 1)
    lab[0] = 1
    lab[-1] = 1 + min(1, n_mat - 1)
    for _ in range(2):
        sx, sy, sz = ri(1, 2), ri(1, 2), ri(1, 2)
        x0, y0, z0 = ri(2, X - 2 - sx), ri(2, Y - 2 - sy), ri(2, Z - 2 - sz)
        lab[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = ri(1, n_mat)
    table = np.zeros((n_mat, 20), np.float32)
    for m in range(n_mat):
        table[m, :] = np.float32(0.01 + 0.6 * u())
    src = [(ri(1, X - 2), ri(1, Y - 2), ri(1, Z - 2)) for _ in range(2)]
    kinds = [ri(0, 1) for _ in range(2)]
    rec = [(ri(1, X - 2), ri(1, Y - 2), ri(1, Z - 2)) for _ in range(3)]
    sig = np.array([[2 * u() - 1 for _ in range(n_steps)] for _ in range(2)])
    pos, mat = classify(lab)
    slot = lambda m: (int(np.float32(m * 20) * np.float32(octave)) if literal & 1 else m * 20 + octave)
    betas = [table.reshape(-1)[slot(m)] if slot(m) < table.size else 0.0 for m in range(256)]
    lam = math.sqrt(1.0 / 3.0)
    r64, P64, Pp64 = run(pos, mat, np.array(betas, np.float32).astype(np.float64), octave, lam,
                         src, kinds, sig, rec, n_steps, np.float64, literal)
    r32, P32, Pp32 = run(pos, mat, np.array(betas, np.float32), octave, lam, src, kinds,
                         sig.astype(np.float32), rec, n_steps, np.float32, literal)
    scale = 9
    rs16, Ps16, Pps16 = run_f16(pos, mat, betas, lam, src, kinds, sig, rec, n_steps, scale, literal)
    rd16, Pd16, Dd16 = run_f16(pos, mat, betas, lam, src, kinds, sig, rec, n_steps, scale, literal,
                               diff=True)
    rsd16, Ssd16, Dsd16 = run_sumdiff16(pos, mat, betas, lam, src, kinds, sig, rec, n_steps, scale,
                                        literal)
    return dict(labels=lab, table=table, octave=octave, literal=literal, lam=lam,
                src=np.array(src, np.int32), kinds=np.array(kinds, np.int32), signals=sig,
                rec=np.array(rec, np.int32), n_steps=n_steps, pos=pos, mat=mat,
                resp64=r64, P64=P64, Pp64=Pp64, resp32=r32, P32=P32, Pp32=Pp32,
                scale16=scale, resp_s16=rs16, P_s16=Ps16, Pp_s16=Pps16, resp_d16=rd16, P_d16=Pd16,
                D_d16=Dd16, resp_sd16=rsd16, S_sd16=Ssd16, D_sd16=Dsd16)


CASES = [
    ("room_a", 101, (8, 7, 6), 2, 40, 0, 0),
    ("room_b", 102, (9, 8, 8), 3, 50, 0, 3),
    ("room_c_literal", 103, (8, 8, 7), 3, 40, 3, 1),
    ("room_d", 104, (10, 6, 7), 1, 60, 0, 0),
]

if __name__ == "__main__":
    for name, seed, dims, n_mat, n_steps, literal, octave in CASES:
        d = case(seed, dims, n_mat, n_steps, literal, octave)
        np.savez_compressed(OUT / f"{name}.npz", **d)
        print(name, dims, "max |resp64|", float(np.abs(d["resp64"]).max()))
