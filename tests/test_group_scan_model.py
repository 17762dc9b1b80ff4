"""CPU checks of the two structures the device paths rest on (no GPU).

1. The spin-string group scan (mhip_spingroup.hip) finds the connected kets
   of a row as three classes: B — same beta string, alpha distance <= 4;
   A — same alpha string, beta distance 2 or 4; C — alpha distance 2 and beta
   distance 2, reached through the groups one single excitation away on the
   side whose neighbour part is shorter. With fixed electron counts these
   must partition {j : popcount(bra ^ ket_j) <= 4}, the set the reference's
   double loop keeps (double_loop.hpp:45-91); checked here against brute
   force on seeded selected-CI lists, with repeated determinants and in
   random order.
2. The ASCI search's dense keys (mhip_asci.hip): the rank of the beta string
   << bits | the rank of the alpha string, ranks taken over all strings one
   or two excitations from a core string, order candidates as bitset_less
   (beta string first) and are equal exactly when the determinants are.
"""
import numpy as np
import pytest

from macis_amd.synthetic import selected_ci_dets


def _popc(x):
    return np.bitwise_count(x).astype(np.int64)


def _singles(s, uni):
    s = int(s)
    holes = [i for i in range(64) if s >> i & 1]
    parts = [i for i in range(64) if (uni >> i & 1) and not (s >> i & 1)]
    return [s ^ (1 << h) ^ (1 << p) for h in holes for p in parts]


def _group_scan_columns(a, b):
    """Columns of every row by the scan's classes (unsorted sets)."""
    n = len(a)
    ga = {}
    gb = {}
    for i in range(n):
        ga.setdefault(int(a[i]), []).append(i)
        gb.setdefault(int(b[i]), []).append(i)
    uni_a = int(np.bitwise_or.reduce(a))
    uni_b = int(np.bitwise_or.reduce(b))
    nb_a = {s: [t for t in _singles(s, uni_a) if t in ga] for s in ga}
    nb_b = {s: [t for t in _singles(s, uni_b) if t in gb] for s in gb}
    size_a = {s: sum(len(ga[t]) for t in nb_a[s]) for s in ga}
    size_b = {s: sum(len(gb[t]) for t in nb_b[s]) for s in gb}
    out = []
    for i in range(n):
        x, y = int(a[i]), int(b[i])
        cols = []
        # class B: the row's beta group, alpha distance <= 4
        cols += [k for k in gb[y] if bin(int(a[k]) ^ x).count("1") <= 4]
        # class A: the row's alpha group, beta distance 2 or 4
        cols += [k for k in ga[x] if bin(int(b[k]) ^ y).count("1") in (2, 4)]
        # class C through the side with the shorter neighbour part
        if size_a[x] <= size_b[y]:
            for t in nb_a[x]:
                cols += [k for k in ga[t] if bin(int(b[k]) ^ y).count("1") == 2]
        else:
            for t in nb_b[y]:
                cols += [k for k in gb[t] if bin(int(a[k]) ^ x).count("1") == 2]
        out.append(cols)
    return out


@pytest.mark.parametrize("case", ["sorted", "shuffled_dups", "wide_core"])
def test_group_scan_classes_partition_connected_set(case):
    rng = np.random.default_rng(21)
    if case == "wide_core":
        d = selected_ci_dets(20, 6, 6, 700, seed=22, energies=0.3 * np.arange(20))
    else:
        d = selected_ci_dets(16, 5, 4, 900, seed=23, energies=0.25 * np.arange(16))
    if case == "shuffled_dups":
        d = np.concatenate([d, d[rng.choice(len(d), 25, replace=False)]])
        d = d[rng.permutation(len(d))]
    a, b = d[:, 0], d[:, 1]
    cols = _group_scan_columns(a, b)
    for i in range(len(d)):
        e = _popc(a ^ a[i]) + _popc(b ^ b[i])
        ref = np.nonzero(e <= 4)[0]
        got = np.array(cols[i], dtype=np.int64)
        assert len(got) == len(np.unique(got)), f"row {i}: a ket reached twice"
        assert np.array_equal(np.sort(got), ref), f"row {i}"


def test_group_scan_needs_fixed_electron_counts():
    # with the alpha count varying, an odd alpha distance plus an odd beta
    # distance can reach 4 in total: such kets fall in no class, which is why
    # the device path refuses these lists
    a = np.array([0b0111, 0b1011, 0b0011], dtype=np.uint64)
    b = np.array([0b0001, 0b0001, 0b0011], dtype=np.uint64)
    e = _popc(a ^ a[0]) + _popc(b ^ b[0])
    assert e[2] == 2 and _popc(a ^ a[0])[2] == 1  # alpha distance 1
    cols = _group_scan_columns(a, b)
    assert 2 not in cols[0]


def test_dense_asci_keys_keep_bitset_order():
    rng = np.random.default_rng(5)
    norb = 14
    core = selected_ci_dets(norb, 4, 4, 12, seed=6, energies=0.3 * np.arange(norb))
    strings = []
    for spin in (0, 1):
        s = set()
        for x in np.unique(core[:, spin]):
            x = int(x)
            s.add(x)
            occ = [i for i in range(norb) if x >> i & 1]
            vir = [i for i in range(norb) if not x >> i & 1]
            for h in occ:
                for p in vir:
                    s.add(x ^ (1 << h) ^ (1 << p))
            for i1 in range(len(occ)):
                for i2 in range(i1 + 1, len(occ)):
                    for v1 in range(len(vir)):
                        for v2 in range(v1 + 1, len(vir)):
                            s.add(x ^ (1 << occ[i1]) ^ (1 << occ[i2]) ^ (1 << vir[v1])
                                  ^ (1 << vir[v2]))
        strings.append(np.array(sorted(s), dtype=np.uint64))
    bits_a = int(np.ceil(np.log2(max(len(strings[0]), 2))))
    # candidates: random pairs of table strings
    ia = rng.integers(0, len(strings[0]), 4000)
    ib = rng.integers(0, len(strings[1]), 4000)
    ia[:100] = ia[100:200]  # repeated determinants
    ib[:100] = ib[100:200]
    alpha, beta = strings[0][ia], strings[1][ib]
    ra = np.searchsorted(strings[0], alpha)
    rb = np.searchsorted(strings[1], beta)
    key = (rb.astype(np.uint64) << np.uint64(bits_a)) | ra.astype(np.uint64)
    # bitset_less: beta string first, then alpha (bitset_operations.hpp:289-307)
    full = [(int(y), int(x)) for x, y in zip(alpha, beta)]
    order_full = sorted(range(len(full)), key=lambda i: (full[i], i))
    order_key = sorted(range(len(full)), key=lambda i: (int(key[i]), i))
    assert order_full == order_key
    for i in range(0, 4000, 37):
        for j in range(0, 4000, 41):
            assert (key[i] == key[j]) == (full[i] == full[j])
