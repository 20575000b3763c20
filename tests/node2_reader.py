"""Independent reader of node2 NanoVDB grids for the test suite.

Follows the node2 layout of nanovdb_editor/PNanoVDBExt.h:28-83 (word offsets, fanouts, packed
16-bit prefix lanes, relative child pointers, keyed root) and the NanoVDB header fields that the
grid builder fills (GridData 672 B, TreeData 64 B, GridBlindMetaData 288 B). It walks the tree
from the root, lists every node and every active voxel with its value index, and checks the
structural invariants the reference grid builder establishes (grid_build_count_new.slang,
grid_build_scatter_new.slang, grid_build_blindmetadata.slang).
"""
import numpy as np

ROOT, UPPER, LOWER, LEAF = 0, 1, 2, 3
OFF_NODE_IDX, OFF_VALUE_IDX = 0, 1
OFF_VMP = (258, 258, 34, 6)
OFF_VM = (386, 386, 50, 8)
OFF_CMP = (898, 898, 114, 6)
OFF_CM = (1026, 1026, 130, 8)
OFF_KMP = (1538, 898, 114, 6)
OFF_KM = (1666, 1026, 130, 8)
OFF_CHILDREN = (2180, 1538, 194, 16)
FANOUT = (32768, 32768, 4096, 512)
FANOUT_BITS = (5, 5, 4, 3)
TILE_BITS = (12, 7, 3, 0)
MAX_SIZE = (541728, 274464, 34336, 128)
END_KEY = 1 << 63
MAGIC = 0x314244566F6E614E  # "NanoVDB1"
GRID_TYPE_NODE2 = 27


def _popcount64(words):
    w = np.asarray(words, dtype=np.uint64)
    b = w.view(np.uint8)
    return np.unpackbits(b).reshape(w.size, 64).sum(axis=1).astype(np.int64)


def _bits(words):
    """Indices of the set bits of a little-endian mask word array."""
    w = np.ascontiguousarray(words, dtype="<u8")
    return np.flatnonzero(np.unpackbits(w.view(np.uint8), bitorder="little"))


class Node2Grid:
    def __init__(self, data):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        pad = (-data.size) % 8
        if pad:
            data = np.concatenate([data, np.zeros(pad, np.uint8)])
        self.bytes = data
        self.w = data.view("<u8")
        assert int(self.w[0]) == MAGIC, "bad magic"
        self.grid_size = int(self.w[32 // 8])
        self.grid_type = int(data[636:640].view("<u4")[0])
        assert self.grid_type == GRID_TYPE_NODE2
        tree = 672
        self.root = (tree + int(self.w[(tree + 24) // 8])) // 8  # word index of the root
        self.count_leaf, self.count_lower, self.count_upper = (int(x) for x in data[tree + 32:tree + 44].view("<u4"))
        self.blind_offset = int(self.w[640 // 8])
        self.blind_count = int(data[648:652].view("<u4")[0])

    # --- raw accessors
    def r(self, idx64):
        return int(self.w[idx64])

    def words(self, idx64, count):
        return self.w[idx64:idx64 + count]

    def blind(self, i):
        """(data byte offset, value count, value size, data type) of blind metadata entry i."""
        meta = self.blind_offset + 288 * i
        off = meta + int(self.w[meta // 8])
        count = int(self.w[meta // 8 + 1])
        size = int(self.bytes[meta + 16:meta + 20].view("<u4")[0])
        dtype = int(self.bytes[meta + 28:meta + 32].view("<u4")[0])
        return off, count, size, dtype

    def values_u32(self):
        off, count, size, _ = self.blind(1)
        assert size == 4
        return self.bytes[off:off + 4 * count].view("<u4")

    def values_f32(self):
        off, count, size, _ = self.blind(1)
        assert size == 4
        return self.bytes[off:off + 4 * count].view("<f4")

    # --- tree walk
    def walk(self):
        """Returns (nodes, voxels): nodes is a list of dicts (addr64, type, origin, node_idx, value_idx,
        value_count); voxels is an (M, 4) int64 array of (i, j, k, value_idx) in tree order."""
        nodes = []
        vox = []
        root = self.root
        nodes.append(dict(addr=root, type=ROOT, origin=None, node_idx=self.r(root + OFF_NODE_IDX),
                          value_idx=self.r(root + OFF_VALUE_IDX),
                          value_count=int(_popcount64(self.words(root + OFF_VM[ROOT], 512)).sum()),
                          parent=None, slot=0))
        cm = self.words(root + OFF_CM[ROOT], 512)
        km = self.words(root + OFF_KM[ROOT], 512)
        child_count = self.r(root + OFF_CMP[ROOT]) & 0xFFFF
        key_slots = _bits(km)
        for rank, n in enumerate(_bits(cm)):
            child = root + self.r(root + OFF_CHILDREN[ROOT] + rank) // 8
            kidx = int(np.searchsorted(key_slots, n))
            assert kidx < key_slots.size and key_slots[kidx] == n, "root child without a key"
            key = self.r(root + OFF_CHILDREN[ROOT] + child_count + kidx)
            hi = [(key >> 42) & 0x1FFFFF, (key >> 21) & 0x1FFFFF, key & 0x1FFFFF]
            origin = [int(np.int32(np.uint32((h << 12) & 0xFFFFFFFF))) for h in hi]
            # the direct slot of the key must be the slot it is stored in (no probing in these grids)
            slot = ((origin[0] & 0x1FFFF) >> 12) << 10 | ((origin[1] & 0x1FFFF) >> 12) << 5 | ((origin[2] & 0x1FFFF) >> 12)
            assert slot == n, "root key stored away from its direct slot"
            self._walk_node(child, UPPER, origin, nodes, vox, root, int(n))
        voxels = np.array(vox, dtype=np.int64).reshape(-1, 4)
        return nodes, voxels

    def _walk_node(self, node, t, origin, nodes, vox, parent, slot):
        nwords = FANOUT[t] // 64
        vm = self.words(node + OFF_VM[t], nwords)
        nodes.append(dict(addr=node, type=t, origin=tuple(origin), node_idx=self.r(node + OFF_NODE_IDX),
                          value_idx=self.r(node + OFF_VALUE_IDX), value_count=int(_popcount64(vm).sum()),
                          parent=parent, slot=slot))
        fb, tb = FANOUT_BITS[t], TILE_BITS[t]
        fmask = (1 << fb) - 1
        if t == LEAF:
            vbase = self.r(node + OFF_VALUE_IDX)
            for rank, n in enumerate(_bits(vm)):
                vox.append((origin[0] + ((n >> 6) & 7), origin[1] + ((n >> 3) & 7), origin[2] + (n & 7), vbase + rank))
            return
        cm = self.words(node + OFF_CM[t], nwords)
        for rank, n in enumerate(_bits(cm)):
            child = node + self.r(node + OFF_CHILDREN[t] + rank) // 8
            o = (origin[0] + (((n >> (2 * fb)) & fmask) << tb), origin[1] + (((n >> fb) & fmask) << tb),
                 origin[2] + ((n & fmask) << tb))
            self._walk_node(child, t + 1, o, nodes, vox, node, int(n))

    # --- invariants
    def check_prefix(self, node, t, prefix_off, mask_off):
        nwords = FANOUT[t] // 64
        mask = self.words(node + mask_off, nwords)
        pc = _popcount64(mask)
        excl = np.concatenate([[0], np.cumsum(pc)[:-1]])
        total = int(pc.sum())
        for q in range((nwords + 3) // 4):
            word = self.r(node + prefix_off + q)
            for lane in range(4):
                wi = 4 * q + lane
                if wi >= nwords:
                    continue
                got = (word >> (16 * lane)) & 0xFFFF
                want = (total if (q == 0 and lane == 0) else int(excl[wi])) & 0xFFFF
                assert got == want, f"prefix lane mismatch node {node} type {t} word {wi}: {got} != {want}"
        return total

    def check_invariants(self):
        """Structural invariants of a node2 grid written by the raster3d builder. Returns (nodes, voxels)."""
        nodes, voxels = self.walk()
        for nd in nodes:
            a, t = nd["addr"], nd["type"]
            self.check_prefix(a, t, OFF_VMP[t], OFF_VM[t])
            if t != LEAF:
                nchild = self.check_prefix(a, t, OFF_CMP[t], OFF_CM[t])
                if t == ROOT:
                    self.check_prefix(a, t, OFF_KMP[t], OFF_KM[t])
                    assert np.array_equal(self.words(a + OFF_KM[t], 512), self.words(a + OFF_CM[t], 512))
                    # the key table follows the children, all keys present, then end keys never appear
                    keys = self.words(a + OFF_CHILDREN[t] + nchild, nchild)
                    assert not np.any(keys == np.uint64(END_KEY))
                # internal nodes of a raster grid carry no tiles
                assert nd["value_count"] == 0 or t == ROOT
        # node counts agree with the tree header
        types = [nd["type"] for nd in nodes]
        assert types.count(UPPER) == self.count_upper
        assert types.count(LOWER) == self.count_lower
        assert types.count(LEAF) == self.count_leaf
        # node_idx is a permutation of 0..count-1 and value_idx is 2 + the exclusive scan of the
        # value counts in node_idx order
        order = sorted(nodes, key=lambda d: d["node_idx"])
        assert [d["node_idx"] for d in order] == list(range(len(nodes)))
        acc = 2
        for d in order:
            assert d["value_idx"] == acc, f"value_idx {d['value_idx']} != {acc} at node_idx {d['node_idx']}"
            acc += d["value_count"]
        # every node lies inside the node segment, at its own size, without overlap
        spans = sorted((d["addr"] * 8, d["addr"] * 8 + MAX_SIZE[d["type"]]) for d in nodes)
        for (a0, a1), (b0, _) in zip(spans, spans[1:]):
            assert a1 <= b0, "overlapping nodes"
        # blind data: [0] bbox, [1] one value per voxel plus the two background values
        assert self.blind_count == 2
        _, vcount, vsize, _ = self.blind(1)
        assert vcount == acc and vsize == 4
        off1 = self.blind(1)[0]
        assert off1 + 4 * vcount <= self.grid_size
        # value indices of the voxels are exactly 2 .. 2 + V - 1
        if voxels.size:
            assert np.array_equal(np.sort(voxels[:, 3]), np.arange(2, 2 + voxels.shape[0]))
        return nodes, voxels

    def bbox(self):
        off, count, size, _ = self.blind(0)
        return self.bytes[off:off + 24].view("<i4").reshape(2, 3)

    def bfs_node_order(self, nodes):
        """Expected allocation order of a single-window build: level by level, children in
        (parent allocation rank, slot) order."""
        by_addr = {d["addr"]: d for d in nodes}
        order = [self.root]
        frontier = [self.root]
        for t in (ROOT, UPPER, LOWER):
            nxt = []
            for a in frontier:
                cm = self.words(a + OFF_CM[t], FANOUT[t] // 64)
                for rank, _ in enumerate(_bits(cm)):
                    nxt.append(a + self.r(a + OFF_CHILDREN[t] + rank) // 8)
            order += nxt
            frontier = nxt
        assert len(order) == len(by_addr)
        return order


# ------------------------------------------------------------------- windowed allocation order
# The reference builds the tree window by window (Raster.cpp:450-500, GridBuild.cpp:361-506): the
# kept fragments of window t insert their voxels level by level (root -> upper, upper -> lower,
# lower -> leaf); the children a level pass finds new are allocated after every node allocated so
# far, in (parent allocation rank, slot) order (grid_build_scatter_new.slang). A node is therefore
# created in the first window that holds a kept voxel inside it, and the allocation order is the
# lexicographic order of (window, level, parent rank, slot); node_idx is the allocation rank and the
# nodes lie in the node segment in that order.
NODE_SPAN = {UPPER: 4096, LOWER: 128, LEAF: 8}


def fragment_windows(bbox, footprint, gid, ijk, prim_batch_size, ijk_batch_size):
    """Window of each kept fragment (gid, ijk): fragment slots count the voxels of each prim's AABB
    (x-major, then y, then z) in prim order within a prim batch, windows split each batch's slots
    into runs of ijk_batch_size and are numbered across batches (batches without fragments have no
    window)."""
    n = footprint.size
    PB, W = int(prim_batch_size), int(ijk_batch_size)
    fp = footprint.astype(np.int64)
    nb = (n + PB - 1) // PB
    excl = np.zeros(n, np.int64)
    win_base = np.zeros(nb + 1, np.int64)
    for b in range(nb):
        seg = fp[b * PB:(b + 1) * PB]
        excl[b * PB:(b + 1) * PB] = np.cumsum(seg) - seg
        win_base[b + 1] = win_base[b] + (int(seg.sum()) + W - 1) // W
    g = gid.astype(np.int64)
    mn = bbox[g, :3].astype(np.int64)
    rng = bbox[g, 3:].astype(np.int64) - mn + 1
    d = ijk.astype(np.int64) - mn
    ri = (d[:, 0] * rng[:, 1] + d[:, 1]) * rng[:, 2] + d[:, 2]
    slot = excl[g] + ri
    return win_base[g // PB] + slot // W


def windowed_allocation_order(nodes, voxel_ijk, voxel_window):
    """Node addresses in the expected allocation order, from each node's first-touch window (the
    smallest window among the kept voxels inside it)."""
    first = {}
    for t in (UPPER, LOWER, LEAF):
        span = NODE_SPAN[t]
        org = (voxel_ijk.astype(np.int64) // span) * span
        key = [tuple(o) for o in org]
        for k, w in zip(key, voxel_window.tolist()):
            if first.get((t, k), w + 1) > w:
                first[(t, k)] = w
    root = [d for d in nodes if d["type"] == ROOT][0]
    rank = {root["addr"]: 0}
    order = [root["addr"]]
    by_level = {t: [d for d in nodes if d["type"] == t] for t in (UPPER, LOWER, LEAF)}
    win = {d["addr"]: first[(d["type"], tuple(int(c) for c in d["origin"]))] for t in by_level for d in by_level[t]}
    for w in sorted(set(win.values())):
        for t in (UPPER, LOWER, LEAF):
            new = [d for d in by_level[t] if win[d["addr"]] == w]
            new.sort(key=lambda d: (rank[d["parent"]], d["slot"]))
            for d in new:
                rank[d["addr"]] = len(order)
                order.append(d["addr"])
    return order


# ----------------------------------------------------------------------------- NanoVDB OnIndex
# Node layouts of a NanoVDB OnIndex grid (RootData / Tile / InternalData / LeafData<ValueOnIndex>
# with 32-byte data alignment): byte offsets within each node.
ONI = dict(root_size=96, tile_size=32, upper_vm=32, upper_cm=4128, upper_table=8256, upper_size=270400,
           lower_vm=32, lower_cm=544, lower_table=1088, lower_size=33856, leaf_vm=16, leaf_table=80, leaf_size=96)
GRID_TYPE_ONINDEX = 20


class OnIndexGrid:
    """Reader of a NanoVDB OnIndex grid: root tiles by key, internal tables of child offsets, leaf
    value offsets with the 9-bit packed prefix sums of the leaf value mask."""

    def __init__(self, data):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        pad = (-data.size) % 8
        if pad:
            data = np.concatenate([data, np.zeros(pad, np.uint8)])
        self.b = data
        assert int(data[:8].view("<u8")[0]) == MAGIC
        assert int(data[636:640].view("<u4")[0]) == GRID_TYPE_ONINDEX
        self.root = 672 + int(data[672 + 24:672 + 32].view("<u8")[0])

    def u64(self, off):
        return int(self.b[off:off + 8].view("<u8")[0])

    def i64(self, off):
        return int(self.b[off:off + 8].view("<i8")[0])

    def u32(self, off):
        return int(self.b[off:off + 4].view("<u4")[0])

    def bbox(self):
        return self.b[self.root:self.root + 24].view("<i4").reshape(2, 3)

    def blind(self):
        meta = self.u64(640)
        assert self.u32(648) == 1
        off = meta + self.u64(meta)
        return off, self.u64(meta + 8), self.u32(meta + 16)

    def voxels(self):
        """(M, 4) int64 array of (i, j, k, value index) over all active leaf voxels."""
        out = []
        r = self.root
        tiles = self.u32(r + 24)
        for t in range(tiles):
            tile = r + ONI["root_size"] + ONI["tile_size"] * t
            key = self.u64(tile)
            child = self.i64(tile + 8)
            if child == 0:
                continue
            hi = [(key >> 42) & 0x1FFFFF, (key >> 21) & 0x1FFFFF, key & 0x1FFFFF]
            origin = [int(np.int32(np.uint32((h << 12) & 0xFFFFFFFF))) for h in hi]
            up = r + child
            cm = self.b[up + ONI["upper_cm"]:up + ONI["upper_cm"] + 4096].view("<u8")
            for un in _bits(cm):
                lo = up + self.i64(up + ONI["upper_table"] + 8 * int(un))
                uo = (origin[0] + (((un >> 10) & 31) << 7), origin[1] + (((un >> 5) & 31) << 7), origin[2] + ((un & 31) << 7))
                lcm = self.b[lo + ONI["lower_cm"]:lo + ONI["lower_cm"] + 512].view("<u8")
                for ln in _bits(lcm):
                    lf = lo + self.i64(lo + ONI["lower_table"] + 8 * int(ln))
                    lo_o = (uo[0] + (((ln >> 8) & 15) << 3), uo[1] + (((ln >> 4) & 15) << 3), uo[2] + ((ln & 15) << 3))
                    vm = self.b[lf + ONI["leaf_vm"]:lf + ONI["leaf_vm"] + 64].view("<u8")
                    first = self.u64(lf + ONI["leaf_table"])
                    psum = self.u64(lf + ONI["leaf_table"] + 8)
                    for n in _bits(vm):
                        w = int(n) >> 6
                        below = int(vm[w]) & ((1 << (int(n) & 63)) - 1)
                        before = ((psum >> (9 * (w - 1))) & 511) if w > 0 else 0
                        idx = first + before + bin(below).count("1")
                        out.append((lo_o[0] + ((n >> 6) & 7), lo_o[1] + ((n >> 3) & 7), lo_o[2] + (n & 7), idx))
        return np.array(out, dtype=np.int64).reshape(-1, 4)
