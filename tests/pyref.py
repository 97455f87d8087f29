"""Independent pure-Python model of the commit path, used to pin the oracle.

It shares no code with oracle/oracle.cpp and is written from the mathematics
rather than from the oracle:
  * Blake2s is Python's hashlib (RFC 7693);
  * QM31 products come from generic polynomial arithmetic in (i, u) reduced by
    i^2 = -1, u^2 = 2 + i (not from the (2p - q, p + 2q) shortcut);
  * circle points are G^e computed from their exponents (G has order 2^31),
    so coset points, NTT twiddles and FRI twiddles follow from exponent
    arithmetic instead of repeated doubling / squashing;
  * the native NTT is evaluated directly as sum_j c_j * b_j(p) with the basis
    product b_j(p) = prod_{bit_b(j)=1} (-1)^{bit_b(p)} T_b[p >> (b+1)], which
    is what the layer network of cuda/circle_ntt.cu computes (derived by
    induction over its layers); O(n^2), tiny sizes only.
"""

from __future__ import annotations

import hashlib
import struct
from functools import lru_cache

P = (1 << 31) - 1
G = (2, 1268011823)


# ------------------------------------------------------------------ field
def inv(a: int) -> int:
    return pow(a, P - 2, P)


def qmul(x, y):
    # x = x0 + x1 i + x2 u + x3 i u ; generic product then reduction.
    terms = {}
    for a, (ea, eb) in zip(x, [(0, 0), (1, 0), (0, 1), (1, 1)]):
        for b, (fa, fb) in zip(y, [(0, 0), (1, 0), (0, 1), (1, 1)]):
            key = (ea + fa, eb + fb)
            terms[key] = (terms.get(key, 0) + a * b) % P
    # reduce powers: i^2 = -1, u^2 = 2 + i
    changed = True
    while changed:
        changed = False
        for (ei, eu), c in list(terms.items()):
            if c == 0:
                del terms[(ei, eu)]
                continue
            if ei >= 2:
                del terms[(ei, eu)]
                k = (ei - 2, eu)
                terms[k] = (terms.get(k, 0) - c) % P
                changed = True
            elif eu >= 2:
                del terms[(ei, eu)]
                for k, f in (((ei, eu - 2), 2), ((ei + 1, eu - 2), 1)):
                    terms[k] = (terms.get(k, 0) + f * c) % P
                changed = True
    return tuple(terms.get(k, 0) for k in [(0, 0), (1, 0), (0, 1), (1, 1)])


def qadd(x, y):
    return tuple((a + b) % P for a, b in zip(x, y))


def qsub(x, y):
    return tuple((a - b) % P for a, b in zip(x, y))


def qscale(x, s):
    return tuple(a * s % P for a in x)


# ------------------------------------------------------------------ circle
def pmul(p, q):
    return ((p[0] * q[0] - p[1] * q[1]) % P, (p[0] * q[1] + p[1] * q[0]) % P)


@lru_cache(maxsize=None)
def gpow(e: int):
    e %= 1 << 31
    r, b = (1, 0), G
    while e:
        if e & 1:
            r = pmul(r, b)
        b = pmul(b, b)
        e >>= 1
    return r


def half_coset_exp(k: int, i: int) -> int:
    # half_coset(k).at(i) = G^(2^(30-k)) * G^(i * 2^(31-k)) = G^((2i+1) 2^(30-k))
    return ((2 * i + 1) << (30 - k)) % (1 << 31)


def half_odds_exp(k: int, i: int) -> int:
    # half_odds(k).at(i) = G^(2^(29-k)) * G^(i * 2^(31-k)) = G^((4i+1) 2^(29-k))   (k <= 29)
    return ((4 * i + 1) << (29 - k)) % (1 << 31)


def bitrev(x: int, bits: int) -> int:
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


# ------------------------------------------------------------------ NTT
def ntt_tables(k: int):
    """T_b tables of a 2^k native NTT on half_coset(k): T_0 = circle (y), T_b = line layer b-1."""
    n = 1 << k
    circle = [gpow(half_coset_exp(k, i))[1] for i in range(n // 2)]
    tables = [circle]
    for lyr in range(k - 1):
        # line layer l entry i = X_l[2i]; X_l[j] = x(initial^(2^l) * step^(j 4^l)) with
        # initial = G^(2^(30-k)), step = G^(2^(31-k)).
        size = n >> (lyr + 1)
        row = []
        for i in range(size):
            j = 2 * i
            e = ((1 << lyr) << (30 - k)) + j * (4**lyr) * (1 << (31 - k))
            row.append(gpow(e)[0])
        tables.append(row)
    return tables


def ntt_eval_closed_form(coeffs, k: int):
    n = 1 << k
    T = ntt_tables(k)
    out = []
    for p in range(n):
        factors = []
        for b in range(k):
            w = T[b][p >> (b + 1)]
            factors.append((P - w) % P if (p >> b) & 1 else w)
        acc = 0
        for j, c in enumerate(coeffs):
            if c == 0:
                continue
            t = c
            for b in range(k):
                if (j >> b) & 1:
                    t = t * factors[b] % P
            acc += t
        out.append(acc % P)
    return out


def ntt_interp_closed_form(values, k: int):
    """Inverse of ntt_eval_closed_form by Gaussian elimination (tiny k only)."""
    n = 1 << k
    T = ntt_tables(k)
    rows = []
    for p in range(n):
        factors = []
        for b in range(k):
            w = T[b][p >> (b + 1)]
            factors.append((P - w) % P if (p >> b) & 1 else w)
        row = []
        for j in range(n):
            t = 1
            for b in range(k):
                if (j >> b) & 1:
                    t = t * factors[b] % P
            row.append(t)
        rows.append(row + [values[p] % P])
    # solve
    for col in range(n):
        piv = next(r for r in range(col, n) if rows[r][col])
        rows[col], rows[piv] = rows[piv], rows[col]
        iv = inv(rows[col][col])
        rows[col] = [v * iv % P for v in rows[col]]
        for r in range(n):
            if r != col and rows[r][col]:
                f = rows[r][col]
                rows[r] = [(a - f * b) % P for a, b in zip(rows[r], rows[col])]
    return [rows[r][n] for r in range(n)]


# ------------------------------------------------------------------ Blake2s / Merkle / channel
def h2w(d: bytes):
    return list(struct.unpack("<8I", d))


def w2b(ws) -> bytes:
    return struct.pack(f"<{len(ws)}I", *ws)


def leaf_hash(words):
    return h2w(hashlib.blake2s(w2b(words)).digest())


def node_hash(l, r):
    return h2w(hashlib.blake2s(w2b(l) + w2b(r)).digest())


def merkle_layers(leaves):
    layers = [leaves]
    while len(layers[-1]) > 1:
        prev = layers[-1]
        layers.append([node_hash(prev[2 * i], prev[2 * i + 1]) for i in range(len(prev) // 2)])
    return layers


def auth_path(layers, idx):
    path = []
    for layer in layers[:-1]:
        path.append(layer[idx ^ 1])
        idx >>= 1
    return path


class Channel:
    def __init__(self):
        self.state = bytes(32)
        self.n_draws = 0

    def mix_digest(self, d):
        self.state = hashlib.blake2s(self.state + w2b(d)).digest()
        self.n_draws = 0

    def mix_felts(self, words):
        self.state = hashlib.blake2s(self.state + w2b(words)).digest()
        self.n_draws = 0

    def squeeze(self) -> bytes:
        out = hashlib.blake2s(self.state + struct.pack("<I", self.n_draws) + b"\x00").digest()
        self.n_draws += 1
        return out

    def draw_felt(self):
        while True:
            ws = struct.unpack("<8I", self.squeeze())
            if all(w < 2 * P for w in ws):
                return tuple(w % P for w in ws[:4])

    def draw_number(self, bound: int) -> int:
        raw = struct.unpack("<Q", self.squeeze()[:8])[0]
        return raw % bound


# ------------------------------------------------------------------ FRI
def fold_twiddles_circle(k: int):
    return [inv(gpow(half_coset_exp(k, bitrev(2 * i, k)))[1]) for i in range(1 << (k - 1))]


def fold_twiddles_line(k: int):
    return [inv(gpow(half_odds_exp(k, bitrev(2 * i, k)))[0]) for i in range(1 << (k - 1))]


def fold(values, alpha, tw):
    out = []
    for i in range(len(values) // 2):
        f0, f1 = values[2 * i], values[2 * i + 1]
        out.append(qadd(qadd(f0, f1), qmul(alpha, qscale(qsub(f0, f1), tw[i]))))
    return out


# ------------------------------------------------------------------ whole proof (tiny sizes)
def prove(a: int, b: int, log_n: int, n_queries: int = 80, log_last: int = 3) -> bytes:
    n = 1 << log_n
    log_e = log_n + 2
    E = 1 << log_e
    trace = [a % P, b % P]
    while len(trace) < n:
        trace.append((trace[-1] + trace[-2]) % P)
    coeffs = ntt_interp_closed_form(trace, log_n)
    ev = ntt_eval_closed_form(coeffs + [0] * (E - n), log_e)
    ttree = merkle_layers([leaf_hash([v]) for v in ev])
    ch = Channel()
    ch.mix_digest(ttree[-1][0])
    alpha = ch.draw_felt()
    q = []
    for i in range(E):
        c = (ev[(i + 2) % E] - ev[(i + 1) % E] - ev[i]) % P
        q.append(qscale(alpha, c))
    qtree = merkle_layers([leaf_hash(list(v)) for v in q])
    ch.mix_digest(qtree[-1][0])
    layers = [fold(q, ch.draw_felt(), fold_twiddles_circle(log_e))]
    trees = [merkle_layers([leaf_hash(list(v)) for v in layers[0]])]
    ch.mix_digest(trees[0][-1][0])
    cur = log_e - 1
    while cur > log_last:
        al = ch.draw_felt()
        layers.append(fold(layers[-1], al, fold_twiddles_line(cur)))
        cur -= 1
        trees.append(merkle_layers([leaf_hash(list(v)) for v in layers[-1]]))
        ch.mix_digest(trees[-1][-1][0])
    last = layers[-1]
    ch.mix_felts([w for v in last for w in v])
    queries = [ch.draw_number(E) for _ in range(n_queries)]

    out = [0x4B52414B, 1, log_n, a % P, b % P] + ttree[-1][0] + qtree[-1][0]
    out += [len(trees)] + [w for t in trees for w in t[-1][0]]
    out += [len(last)] + [w for v in last for w in v]
    out += [len(queries)] + queries

    def decom(tree, vals, idxs, width):
        res = [len(idxs)]
        for i in idxs:
            res += list(vals[i]) if width == 4 else [vals[i]]
        for i in idxs:
            res += list(vals[i ^ 1]) if width == 4 else [vals[i ^ 1]]
        for sib in (0, 1):
            for i in idxs:
                pth = auth_path(tree, i ^ sib)
                res.append(len(pth))
                for h in pth:
                    res += h
        return res

    out += decom(ttree, ev, queries, 1)
    out += decom(qtree, q, queries, 4)
    out.append(len(trees))
    fi = [x // 2 for x in queries]
    for t, lay in zip(trees, layers):
        out += decom(t, lay, fi, 4)
        fi = [x // 2 for x in fi]
    return w2b(out)


# ------------------------------------------------------------------ stwo basis (tiny sizes)
def stwo_twiddles(log_n: int):
    """Flat stwo twiddle buffer of a size-2^log_n domain, from exponents: layer j holds
    x(half_odds(log_n-1-j).at(bitrev(i))) for i < 2^(log_n-2-j), then a final 1."""
    out = []
    for j in range(log_n - 1):
        k = log_n - 1 - j
        m = k - 1
        out += [gpow(half_odds_exp(k, bitrev(i, m)))[0] for i in range(1 << m)]
    out.append(1)
    return out


def stwo_point(log_n: int, k: int):
    """Canonic circle-domain point k (natural order) of size 2^log_n."""
    half = 1 << (log_n - 1)
    if k < half:
        return gpow(half_odds_exp(log_n - 1, k))
    x, y = gpow(half_odds_exp(log_n - 1, k - half))
    return (x, (P - y) % P)


def eval_circle_poly(coeffs, point):
    """stwo circle-polynomial evaluation by recursive folding (src/oods.rs:133-215):
    factors [pi^(k-2)(x), ..., pi(x), x, y] for the halves from outermost to innermost."""
    k = len(coeffs).bit_length() - 1
    x, y = point
    raw = [y]
    for _ in range(1, k):
        raw.append(x)
        x = (2 * x * x - 1) % P
    factors = raw[::-1]

    def fold(vals, fs):
        if len(vals) == 1:
            return vals[0] % P
        h = len(vals) // 2
        return (fold(vals[:h], fs[1:]) + fs[0] * fold(vals[h:], fs[1:])) % P

    return fold(list(coeffs), factors)


def stwo_evaluate(coeffs, log_n: int):
    """Evaluations at the canonic domain in bit-reversed order (what the stwo NTT outputs)."""
    return [eval_circle_poly(coeffs, stwo_point(log_n, bitrev(j, log_n))) for j in range(1 << log_n)]


def stwo_interpolate(values, log_n: int):
    n = 1 << log_n
    rows = []
    for j in range(n):
        pt = stwo_point(log_n, bitrev(j, log_n))
        basis = [eval_circle_poly([1 if t == i else 0 for t in range(n)], pt) for i in range(n)]
        rows.append(basis + [values[j] % P])
    for col in range(n):
        piv = next(r for r in range(col, n) if rows[r][col])
        rows[col], rows[piv] = rows[piv], rows[col]
        iv = inv(rows[col][col])
        rows[col] = [v * iv % P for v in rows[col]]
        for r in range(n):
            if r != col and rows[r][col]:
                f = rows[r][col]
                rows[r] = [(a - f * b) % P for a, b in zip(rows[r], rows[col])]
    return [rows[r][n] for r in range(n)]


def lifted_row(leaf: int, S: int, L: int) -> int:
    if L == S:
        return leaf
    sh = S - L + 1
    return ((leaf >> sh) << 1) | (leaf & 1)


def lifted_leaf(cols, logs, S, leaf):
    return h2w(hashlib.blake2s(w2b([cols[c][lifted_row(leaf, S, logs[c])] for c in range(len(cols))])).digest())


def wide_commit(trace_cols, logs, blowup=2, log_last=2):
    """Batched wide commitment (tiny sizes): stwo LDE of each column, lifted tree, then FRI
    on f = sum_c beta^c lifted_c with the canonic-domain circle fold and half_odds line folds."""
    ev = []
    for col, l in zip(trace_cols, logs):
        coef = stwo_interpolate(col, l)
        ev.append(stwo_evaluate(coef + [0] * ((1 << (l + blowup)) - (1 << l)), l + blowup))
    elogs = [l + blowup for l in logs]
    S = max(elogs)
    n = 1 << S
    leaves = [lifted_leaf(ev, elogs, S, i) for i in range(n)]
    root = merkle_layers(leaves)[-1][0]
    ch = Channel()
    ch.mix_digest(root)
    beta = ch.draw_felt()
    pw = [(1, 0, 0, 0)]
    for _ in range(len(ev) - 1):
        pw.append(qmul(pw[-1], beta))
    f = []
    for i in range(n):
        acc = (0, 0, 0, 0)
        for c in range(len(ev)):
            acc = qadd(acc, qscale(pw[c], ev[c][lifted_row(i, S, elogs[c])]))
        f.append(acc)
    tw = [inv(gpow(half_odds_exp(S - 1, bitrev(i, S - 1)))[1]) for i in range(n // 2)]
    a0 = ch.draw_felt()
    layers = [fold(f, a0, tw)]
    roots = []
    t = merkle_layers([leaf_hash(list(v)) for v in layers[0]])
    roots.append(t[-1][0])
    ch.mix_digest(roots[-1])
    cur = S - 1
    while cur > log_last:
        al = ch.draw_felt()
        layers.append(fold(layers[-1], al, fold_twiddles_line(cur)))
        cur -= 1
        t = merkle_layers([leaf_hash(list(v)) for v in layers[-1]])
        roots.append(t[-1][0])
        ch.mix_digest(roots[-1])
    return root, roots, [w for v in layers[-1] for w in v], beta
