"""Vectorised reader and invariant check of node2 NanoVDB grids, for grids of millions of leaves.

Same layout and the same invariants as node2_reader.Node2Grid (nanovdb_editor/PNanoVDBExt.h:28-83;
grid_build_count_new.slang, grid_build_scatter_new.slang, grid_build_blindmetadata.slang), but
level by level with numpy instead of a per-voxel walk: child lists come from the set bits of all
masks of a level at once, prefix lanes are compared as arrays, and the voxel bounding box is
derived from the leaf masks (word index = x, byte = y, bit = z) without listing voxels. The CPU
tests check that both readers agree on the same grids.
"""
import numpy as np

from node2_reader import (END_KEY, FANOUT, FANOUT_BITS, GRID_TYPE_NODE2, LEAF, LOWER, MAGIC, MAX_SIZE, OFF_CHILDREN,
                          OFF_CM, OFF_CMP, OFF_KM, OFF_KMP, OFF_NODE_IDX, OFF_VALUE_IDX, OFF_VM, OFF_VMP, ROOT,
                          TILE_BITS, UPPER)


def _set_bits_2d(masks):
    """(node row, bit index) of every set bit of a (K, nwords) little-endian u64 mask array, in
    row-major order (so bits of one node appear in increasing index order)."""
    m = np.ascontiguousarray(masks, dtype="<u8")
    if m.size == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    rows = []
    cols = []
    # chunk the rows so the unpacked bits stay below ~256 MB
    per = max(1, (1 << 28) // (m.shape[1] * 64))
    for r0 in range(0, m.shape[0], per):
        b = np.unpackbits(m[r0:r0 + per].view(np.uint8), axis=1, bitorder="little")
        r, c = np.nonzero(b)
        rows.append(r.astype(np.int64) + r0)
        cols.append(c.astype(np.int64))
    return np.concatenate(rows), np.concatenate(cols)


def _ranks(rows, count):
    """Rank of each entry among the entries of its row (rows sorted ascending)."""
    first = np.searchsorted(rows, np.arange(count))
    return np.arange(rows.size, dtype=np.int64) - first[rows]


class FastNode2:
    def __init__(self, data):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        pad = (-data.size) % 8
        if pad:
            data = np.concatenate([data, np.zeros(pad, np.uint8)])
        self.bytes = data
        self.w = data.view("<u8")
        assert int(self.w[0]) == MAGIC, "bad magic"
        assert int(data[636:640].view("<u4")[0]) == GRID_TYPE_NODE2
        self.grid_size = int(self.w[4])
        tree = 672
        self.root = (tree + int(self.w[(tree + 24) // 8])) // 8
        self.count_leaf, self.count_lower, self.count_upper = (int(x) for x in data[tree + 32:tree + 44].view("<u4"))
        self.blind_offset = int(self.w[640 // 8])
        self.blind_count = int(data[648:652].view("<u4")[0])
        self._levels = None

    def _gather(self, addr, off, count):
        idx = addr[:, None] + off + np.arange(count, dtype=np.int64)[None, :]
        return self.w[idx]

    def blind(self, i):
        meta = self.blind_offset + 288 * i
        off = meta + int(self.w[meta // 8])
        count = int(self.w[meta // 8 + 1])
        size = int(self.bytes[meta + 16:meta + 20].view("<u4")[0])
        return off, count, size

    def bbox(self):
        off, _, _ = self.blind(0)
        return self.bytes[off:off + 24].view("<i4").reshape(2, 3)

    def levels(self):
        """Per level: dict(addr (K,) word index, origin (K, 3), parent (K,) row of the parent level)."""
        if self._levels is not None:
            return self._levels
        w = self.w
        root = self.root
        cm = w[root + OFF_CM[ROOT]:root + OFF_CM[ROOT] + 512]
        km = w[root + OFF_KM[ROOT]:root + OFF_KM[ROOT] + 512]
        assert np.array_equal(cm, km), "root key mask differs from its child mask"
        _, slots = _set_bits_2d(cm[None, :])
        nchild = slots.size
        offs = w[root + OFF_CHILDREN[ROOT]:root + OFF_CHILDREN[ROOT] + nchild].astype(np.int64)
        keys = w[root + OFF_CHILDREN[ROOT] + nchild:root + OFF_CHILDREN[ROOT] + 2 * nchild]
        assert not np.any(keys == np.uint64(END_KEY)), "end key inside the key table"
        hi = np.stack([(keys >> np.uint64(42)) & np.uint64(0x1FFFFF), (keys >> np.uint64(21)) & np.uint64(0x1FFFFF),
                       keys & np.uint64(0x1FFFFF)], 1)
        origin = ((hi << np.uint64(12)) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32).astype(np.int64)
        direct = (((origin[:, 0] & 0x1FFFF) >> 12) << 10) | (((origin[:, 1] & 0x1FFFF) >> 12) << 5) | \
                 ((origin[:, 2] & 0x1FFFF) >> 12)
        assert np.array_equal(direct, slots), "root key stored away from its direct slot"
        levels = {ROOT: dict(addr=np.array([root], np.int64), origin=np.zeros((1, 3), np.int64),
                             parent=np.zeros(1, np.int64))}
        levels[UPPER] = dict(addr=root + offs // 8, origin=origin, parent=np.zeros(nchild, np.int64))
        for t in (UPPER, LOWER):
            A, O = levels[t]["addr"], levels[t]["origin"]
            nw = FANOUT[t] // 64
            rows, bits = _set_bits_2d(self._gather(A, OFF_CM[t], nw))
            rank = _ranks(rows, A.size)
            child_off = w[A[rows] + OFF_CHILDREN[t] + rank].astype(np.int64)
            fb, tb = FANOUT_BITS[t], TILE_BITS[t]
            fm = (1 << fb) - 1
            co = O[rows] + np.stack([((bits >> (2 * fb)) & fm) << tb, ((bits >> fb) & fm) << tb, (bits & fm) << tb], 1)
            levels[t + 1] = dict(addr=A[rows] + child_off // 8, origin=co, parent=rows)
        self._levels = levels
        return levels

    def _check_prefix(self, A, t, prefix_off, mask_off):
        nw = FANOUT[t] // 64
        M = self._gather(A, mask_off, nw)
        pc = np.bitwise_count(M).astype(np.int64)
        excl = np.cumsum(pc, axis=1) - pc
        total = pc.sum(axis=1)
        P = self._gather(A, prefix_off, (nw + 3) // 4)
        lanes = np.stack([(P >> np.uint64(16 * l)) & np.uint64(0xFFFF) for l in range(4)], 2).reshape(A.size, -1)[:, :nw]
        want = (excl & 0xFFFF).astype(np.uint64)
        want[:, 0] = (total & 0xFFFF).astype(np.uint64)
        bad = np.nonzero(lanes != want)
        assert bad[0].size == 0, f"prefix lane mismatch: type {t}, {bad[0].size} lanes, first node {A[bad[0][0]]}"
        return total

    def check_invariants(self):
        """Structural invariants of a raster3d node2 grid; returns a summary dict (node counts,
        voxel count, voxel bounding box)."""
        lv = self.levels()
        value_counts = {}
        for t in (ROOT, UPPER, LOWER, LEAF):
            A = lv[t]["addr"]
            value_counts[t] = self._check_prefix(A, t, OFF_VMP[t], OFF_VM[t])
            if t != LEAF:
                nchild = self._check_prefix(A, t, OFF_CMP[t], OFF_CM[t])
                if t == ROOT:
                    self._check_prefix(A, t, OFF_KMP[t], OFF_KM[t])
                # every child mask bit leads to a node, and no internal node is empty
                assert np.all(nchild > 0) or (t == ROOT and A.size == 1)
                assert np.all(value_counts[t] == 0), "internal node with active tiles"
            else:
                assert np.all(value_counts[t] > 0), "leaf without active voxels"
            # children of one level are distinct nodes
            assert np.unique(A).size == A.size
        assert lv[UPPER]["addr"].size == self.count_upper
        assert lv[LOWER]["addr"].size == self.count_lower
        assert lv[LEAF]["addr"].size == self.count_leaf
        # node_idx is a permutation; value_idx = 2 + exclusive scan of value counts in node_idx order
        A = np.concatenate([lv[t]["addr"] for t in (ROOT, UPPER, LOWER, LEAF)])
        vc = np.concatenate([value_counts[t] for t in (ROOT, UPPER, LOWER, LEAF)])
        types = np.concatenate([np.full(lv[t]["addr"].size, t) for t in (ROOT, UPPER, LOWER, LEAF)])
        node_idx = self.w[A + OFF_NODE_IDX].astype(np.int64)
        value_idx = self.w[A + OFF_VALUE_IDX].astype(np.int64)
        order = np.argsort(node_idx, kind="stable")
        assert np.array_equal(node_idx[order], np.arange(A.size)), "node_idx is not a permutation"
        expect = 2 + np.concatenate([[0], np.cumsum(vc[order])[:-1]])
        assert np.array_equal(value_idx[order], expect), "value_idx is not 2 + the exclusive scan"
        V = int(vc.sum())
        # nodes at their own sizes without overlap
        o2 = np.argsort(A)
        start = A[o2] * 8
        end = start + np.array(MAX_SIZE)[types[o2]]
        assert np.all(end[:-1] <= start[1:]), "overlapping nodes"
        # blind data: [0] bbox, [1] one value per voxel plus the two background values
        assert self.blind_count == 2
        off1, vcount, vsize = self.blind(1)
        assert vcount == V + 2 and vsize == 4
        assert off1 + 4 * vcount <= self.grid_size
        # voxel bounding box from the leaf masks: bit n of a leaf is voxel (n >> 6, (n >> 3) & 7, n & 7)
        VM = self._gather(lv[LEAF]["addr"], OFF_VM[LEAF], 8)
        O = lv[LEAF]["origin"]
        lo = np.full(3, np.iinfo(np.int64).max)
        hi = np.full(3, np.iinfo(np.int64).min)
        if VM.shape[0]:
            nzw = VM != 0                                       # x: which words are set
            bytes_ = VM.view(np.uint8).reshape(-1, 8, 8)
            anyb = np.bitwise_or.reduce(bytes_, axis=1)         # y: which bytes of any word
            anybit = np.bitwise_or.reduce(anyb, axis=1)         # z: which bits of any byte
            ar = np.arange(8)
            ybits = anyb != 0
            zbits = np.unpackbits(anybit[:, None], axis=1, bitorder="little").astype(bool)
            for a, m in ((0, nzw), (1, ybits), (2, zbits)):
                lo[a] = (O[:, a] + np.where(m, ar, 8).min(1)).min()
                hi[a] = (O[:, a] + np.where(m, ar, -1).max(1)).max()
        return dict(upper=self.count_upper, lower=self.count_lower, leaf=self.count_leaf, voxels=V,
                    bbox=np.stack([lo, hi]))

    def voxel_ijk_in_box(self, lo, hi):
        """(M, 4) int64 (i, j, k, value index) of the active voxels with lo <= ijk <= hi."""
        lv = self.levels()
        O = lv[LEAF]["origin"]
        sel = np.nonzero(np.all(O + 7 >= lo, 1) & np.all(O <= hi, 1))[0]
        if sel.size == 0:
            return np.zeros((0, 4), np.int64)
        A = lv[LEAF]["addr"][sel]
        rows, bits = _set_bits_2d(self._gather(A, OFF_VM[LEAF], 8))
        rank = _ranks(rows, sel.size)
        ijk = O[sel][rows] + np.stack([bits >> 6, (bits >> 3) & 7, bits & 7], 1)
        vidx = self.w[A[rows] + OFF_VALUE_IDX].astype(np.int64) + rank
        keep = np.all(ijk >= lo, 1) & np.all(ijk <= hi, 1)
        return np.concatenate([ijk[keep], vidx[keep, None]], 1)
