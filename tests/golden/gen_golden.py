"""Writes tests/golden/golden_small.json: known-answer vectors for a tiny
parameter set, computed with Python big integers straight from the defining
formulas (no oracle code, no reference code):

  * primes: q = 1 mod 2n, descending from floor(2^bits/2n)*2n+1 (the rule of
    the reference's src/params.cpp:29-39), primality by trial division;
  * roots: first g^((q-1)/2n) with order exactly 2n (src/params.cpp:42-48);
  * forward NTT: X[k] = sum_j a_j psi^((2 brv(k)+1) j) mod q  (bit-reversed
    evaluation order of the reference's CT transform);
  * basis conversion: (sum_i [x_i (Q/q_i)^-1]_{q_i} (Q/q_i)) mod p_j;
  * keyswitch constants (definitions of src/ks_constants.cpp:31-102);
  * rescale: coefficient form y_i = ((X - centred(X mod q_l)) / q_l) mod q_i,
    then converted to evaluation form with the same brute-force NTT.

Run:  python tests/golden/gen_golden.py
"""
import json
import os
import random

N = 16
BITS_Q0, BITS_Q, BITS_P = 30, 25, 30
L, K, DNUM = 4, 2, 2


def is_prime(x):
    if x < 2:
        return False
    i = 2
    while i * i <= x:
        if x % i == 0:
            return False
        i += 1
    return True


def primes(count, bits, n, avoid=()):
    M = 2 * n
    q = ((1 << bits) // M) * M + 1
    out = []
    while len(out) < count:
        if q not in avoid and is_prime(q):
            out.append(q)
        q -= M
    return out


def root(n, q):
    for g in range(2, q):
        c = pow(g, (q - 1) // (2 * n), q)
        if pow(c, n, q) == q - 1:
            return c
    raise ValueError


def brv(v, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (v & 1)
        v >>= 1
    return r


def ntt_dft(a, q, psi):
    n = len(a)
    lb = n.bit_length() - 1
    return [sum(a[j] * pow(psi, (2 * brv(k, lb) + 1) * j, q) for j in range(n)) % q for k in range(n)]


def prod(xs):
    r = 1
    for x in xs:
        r *= x
    return r


def approx_switch(src_cols, qs, ps):
    """src_cols[i][r] residues mod qs[i] -> dst[j][r] mod ps[j]."""
    Q = prod(qs)
    n = len(src_cols[0])
    out = [[0] * n for _ in ps]
    for r in range(n):
        s = 0
        for i, q in enumerate(qs):
            qh = Q // q
            s += ((src_cols[i][r] * pow(qh, -1, q)) % q) * qh
        for j, p in enumerate(ps):
            out[j][r] = s % p
    return out


def main():
    rnd = random.Random(20260729)
    q = primes(1, BITS_Q0, N) + primes(L - 1, BITS_Q, N)
    p = primes(K, BITS_P, N, set(q))
    roots = [root(N, m) for m in q + p]
    alpha = -(-L // DNUM)

    vec = [rnd.randrange(q[0]) for _ in range(N)]
    ntt_out = ntt_dft(vec, q[0], roots[0])

    # keyswitch constants, definitional
    P = prod(p)
    ksc = {"pInvModq": [pow(P % m, -1, m) for m in q],
           "pHatInv": [pow((P // pj) % pj, -1, pj) for pj in p],
           "pHatModq": [[(P // pj) % m for m in q] for pj in p],
           "parts": []}
    for part in range(DNUM):
        st = alpha * part
        sp = min(alpha, L - st)
        src = q[st:st + sp]
        Qp = prod(src)
        compl = [m for i, m in enumerate(q) if i < st or i >= st + sp] + p
        ksc["parts"].append({
            "startIdx": st, "sizePart": sp, "compl": compl,
            "QHatInv": [pow((Qp // m) % m, -1, m) for m in src],
            "QHatModp": [[(Qp // m) % c for c in compl] for m in src],
        })

    # basis conversion of a random coefficient-form part-0 slice
    part0 = ksc["parts"][0]
    src_cols = [[rnd.randrange(m) for _ in range(N)] for m in q[:part0["sizePart"]]]
    conv = approx_switch(src_cols, q[:part0["sizePart"]], part0["compl"])

    # rescale of a random coefficient-form poly over all L towers
    X = [rnd.randrange(prod(q)) for _ in range(N)]
    coef = [[x % m for x in X] for m in q]
    ql = q[L - 1]
    y = []
    for i in range(L - 1):
        col = []
        for x in X:
            xl = x % ql
            xt = xl - ql if xl > ql // 2 else xl
            col.append(((x - xt) // ql) % q[i])
        y.append(col)
    rescale_in_eval = [ntt_dft(coef[t], q[t], roots[t]) for t in range(L)]
    rescale_out_eval = [ntt_dft(y[t], q[t], roots[t]) for t in range(L - 1)]

    out = dict(n=N, q=q, p=p, roots=roots, dnum=DNUM, alpha=alpha,
               ntt_in=vec, ntt_out=ntt_out, ksc=ksc,
               conv_src=src_cols, conv_out=conv,
               rescale_in_eval=rescale_in_eval, rescale_out_eval=rescale_out_eval)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_small.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
