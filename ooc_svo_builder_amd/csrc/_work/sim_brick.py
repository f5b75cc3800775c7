"""CPU model of the brick fast path (k_brick_level + level formula + k_emit_bricks) vs the oracle."""
import sys
sys.path.insert(0, '/root/repo')
import numpy as np
import oracle
from ooc_svo_builder_amd.formats import NODE_DTYPE

def popc(x): return bin(int(x)).count('1')
def offsets_of(m):
    r = 0; o = []
    for k in range(8):
        if m >> k & 1: o.append(r); r += 1
        else: o.append(-1)
    return o

def build(codes, G, payload):
    D = G.bit_length() - 1
    codes = sorted(int(c) for c in codes)
    V = len(codes)
    # bricks
    bricks = {}
    for c in codes:
        bricks.setdefault(c >> 12, []).append(c)
    bl = sorted(bricks)
    words = {}
    for b in bl:
        w = [0] * 64
        for c in bricks[b]:
            r = c & 4095
            w[r >> 6] |= 1 << (r & 63)
        words[b] = w
    # voxel offsets of bricks
    bvoff = {}
    acc = 0
    for b in bl:
        bvoff[b] = acc; acc += len(bricks[b])
    def shape(w):
        m2 = [sum(1 << j for j in range(8) if (x >> (8 * j)) & 0xFF) for x in w]
        s2 = [(1 + popc(m2[l]) + popc(w[l])) if w[l] else 0 for l in range(64)]
        mask3 = [sum(1 << q for q in range(8) if w[8 * g + q]) for g in range(8)]
        s3 = [(1 + sum(s2[8 * g:8 * g + 8])) if mask3[g] else 0 for g in range(8)]
        mask4 = sum(1 << g for g in range(8) if mask3[g])
        s4 = 1 + sum(s3)
        return m2, s2, mask3, s3, mask4, s4
    # levels >= 4: by prefix (height h node code = c >> 3h)
    lev = {4: [(b, shape(words[b])[5], shape(words[b])[4]) for b in bl]}  # (code,size,mask)
    for h in range(5, D + 1):
        prev = lev[h - 1]
        groups = {}
        for (c, sz, m) in prev:
            groups.setdefault(c >> 3, []).append((c, sz))
        lev[h] = []
        for p in sorted(groups):
            ch = groups[p]
            lev[h].append((p, 1 + sum(s for _, s in ch), sum(1 << (c & 7) for c, _ in ch)))
    nn = lev[D][0][1]
    nodes = [None] * nn
    start = {}
    # top-down heights D..4
    for h in range(D, 3, -1):
        for i, (c, sz, m) in enumerate(lev[h]):
            if h == D:
                st, pos = 0, sz - 1
            else:
                p = c >> 3
                pst = start[(h + 1, p)]
                psz, pm = [(s, mm) for (cc, s, mm) in lev[h + 1] if cc == p][0]
                sib = [(cc, s) for (cc, s, mm) in lev[h] if cc >> 3 == p]
                st = pst + sum(s - 1 for cc, s in sib if cc < c)
                cb = pst + psz - 1 - popc(pm)
                pos = cb + sum(1 for cc, s in sib if cc < c)
            start[(h, c)] = st
            nodes[pos] = (0, st + sz - 1 - popc(m), offsets_of(m))
    # in-brick
    for b in bl:
        w = words[b]
        m2, s2, mask3, s3, mask4, s4 = shape(w)
        S = start[(4, b)]
        cb4 = s4 - 1 - popc(mask4)
        pre3 = 0
        vox = bvoff[b]
        for g in range(8):
            if not mask3[g]: continue
            start3 = S + pre3
            cb3 = start3 + s3[g] - 1 - popc(mask3[g])
            nodes[S + cb4 + popc(mask4 & ((1 << g) - 1))] = (0, cb3, offsets_of(mask3[g]))
            pre2 = 0
            for q in range(8):
                l = 8 * g + q
                if not w[l]: continue
                start2 = start3 + pre2
                cb2 = start2 + s2[l] - 1 - popc(m2[l])
                nodes[cb3 + popc(mask3[g] & ((1 << q) - 1))] = (0, cb2, offsets_of(m2[l]))
                start1 = start2; rank1 = 0
                for j in range(8):
                    byte = (w[l] >> (8 * j)) & 0xFF
                    if not byte: continue
                    nodes[cb2 + rank1] = (0, start1, offsets_of(byte))
                    for r in range(popc(byte)):
                        nodes[start1 + r] = ((1 + vox) if payload else 1, 0, [-1] * 8)
                        vox += 1
                    start1 += popc(byte); rank1 += 1
                pre2 += s2[l] - 1
            pre3 += s3[g] - 1
    arr = np.zeros(nn, NODE_DTYPE)
    for i, n in enumerate(nodes):
        assert n is not None, i
        arr[i]['data'], arr[i]['children_base'], arr[i]['offsets'] = n[0], n[1], n[2]
    return arr

rng = np.random.default_rng(1)
for G in (16, 32, 64):
    for n in (1, 5, 300, 3000):
        for t in range(3):
            codes = np.unique(rng.choice(G ** 3, size=min(n, G**3), replace=False)).astype(np.uint64)
            # clustered too
            if t == 2:
                base = int(rng.integers(0, G**3 - 5000)) if G**3 > 5000 else 0
                codes = np.unique(np.concatenate([codes, np.arange(base, base + min(4096, G**3 - base), dtype=np.uint64)]))
            want, _, _ = oracle.build_octree(codes, G, False)
            got = build(codes, G, False)
            assert got.tobytes() == want.tobytes(), (G, n, t)
print("brick fast-path model matches the oracle")
