"""Host-side M31 / CM31 / QM31 scalar arithmetic and circle-point helpers for the protocol
drivers (channel-level values, a handful per proof). Device data never passes through here.

Field definitions: crates/stwo/src/core/fields/{m31,cm31,qm31}.rs (u^2 = 2 + i, i^2 = -1;
QM31 words [a.0, a.1, b.0, b.1]); circle group: crates/stwo/src/core/circle.rs.
"""
P = (1 << 31) - 1


def minv(a):
    return pow(int(a), P - 2, P)


def _cmul(a0, a1, b0, b1):
    return (a0 * b0 - a1 * b1) % P, (a0 * b1 + a1 * b0) % P


def qadd(x, y):
    return [(int(a) + int(b)) % P for a, b in zip(x, y)]


def qsub(x, y):
    return [(int(a) - int(b)) % P for a, b in zip(x, y)]


def qmul(x, y):
    a0, a1, b0, b1 = (int(v) for v in x)
    c0, c1, d0, d1 = (int(v) for v in y)
    ac = _cmul(a0, a1, c0, c1)
    bd = _cmul(b0, b1, d0, d1)
    rbd = ((2 * bd[0] - bd[1]) % P, (bd[0] + 2 * bd[1]) % P)  # (2 + i) * bd
    ad = _cmul(a0, a1, d0, d1)
    bc = _cmul(b0, b1, c0, c1)
    return [(ac[0] + rbd[0]) % P, (ac[1] + rbd[1]) % P, (ad[0] + bc[0]) % P, (ad[1] + bc[1]) % P]


def qmulm(x, m):
    return [int(v) * int(m) % P for v in x]


def qinv(x):
    a0, a1, b0, b1 = (int(v) for v in x)
    b2 = _cmul(b0, b1, b0, b1)
    ib2 = ((-b2[1]) % P, b2[0])
    a2 = _cmul(a0, a1, a0, a1)
    den = ((a2[0] - 2 * b2[0] - ib2[0]) % P, (a2[1] - 2 * b2[1] - ib2[1]) % P)
    n = minv((den[0] * den[0] + den[1] * den[1]) % P)
    di = (den[0] * n % P, (-den[1]) * n % P)
    x0 = _cmul(a0, a1, di[0], di[1])
    x1 = _cmul(b0, b1, di[0], di[1])
    return [x0[0], x0[1], (-x1[0]) % P, (-x1[1]) % P]


def qpow(x, e):
    r, b = [1, 0, 0, 0], list(x)
    while e:
        if e & 1:
            r = qmul(r, b)
        b = qmul(b, b)
        e >>= 1
    return r


def secure_powers(felt, n):
    """AccumulationOps::generate_secure_powers (crates/stwo/src/prover/air/accumulation.rs:188-196)."""
    out, acc = [], [1, 0, 0, 0]
    for _ in range(n):
        out.append(acc)
        acc = qmul(acc, felt)
    return out


# ---- circle group over M31, by index (order 2^31, generator (2, 1268011823)) ------------------
GEN = (2, 1268011823)


def _padd(p, q):
    return ((p[0] * q[0] - p[1] * q[1]) % P, (p[0] * q[1] + p[1] * q[0]) % P)


def index_to_point(idx):
    idx %= 1 << 31
    r, g = (1, 0), GEN
    while idx:
        if idx & 1:
            r = _padd(r, g)
        g = _padd(g, g)
        idx >>= 1
    return r


def qpoint_double(pt):
    """CirclePoint::double over QM31: (2x^2 - 1, 2xy); points are 8 words (x then y)."""
    x, y = list(pt[:4]), list(pt[4:])
    xx = qmul(x, x)
    return qsub(qadd(xx, xx), [1, 0, 0, 0]) + qmulm(qmul(x, y), 2)


def qpoint_repeated_double(pt, n):
    for _ in range(n):
        pt = qpoint_double(pt)
    return pt


def qpoint_add_m31(pt, q):
    """QM31 point + M31 point (q embedded)."""
    x, y = list(pt[:4]), list(pt[4:])
    return qsub(qmulm(x, q[0]), qmulm(y, q[1])) + qadd(qmulm(x, q[1]), qmulm(y, q[0]))


def double_x(x):
    xx = qmul(x, x)
    return qsub(qadd(xx, xx), [1, 0, 0, 0])


def coset_vanishing_q(coset_initial, coset_log, pt):
    """coset_vanishing (crates/stwo/src/core/constraints.rs:12-35) at a QM31 point:
    x of (p - initial + step/2), doubled coset_log - 1 times."""
    shift = (-coset_initial + ((1 << (31 - coset_log)) >> 1)) % (1 << 31)
    x = qpoint_add_m31(pt, index_to_point(shift))[:4]
    for _ in range(1, coset_log):
        x = double_x(x)
    return x
