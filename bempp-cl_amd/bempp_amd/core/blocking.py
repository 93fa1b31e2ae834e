"""Locality-ordered element blocks with per-block dof slots for the tiled regular kernel.

For spaces whose dofs are shared between elements (P1, RWG/SNC), the regular kernel sums
a workgroup's contributions in an LDS tile indexed by per-block dof slots and adds the
tile to the dense matrix once per entry. This module builds the host side of that
scheme: a space-filling-curve order of the support elements (so that consecutive elements
share dofs), greedy blocks of at most `max_elements` elements and `max_dofs` distinct
dofs, the distinct dof list of each block and the slot of every (element, shape function).
"""

import numpy as _np

# limits of the device tile (BEM_TILE_* in include/bempp_hip.h)
TEST_TILE_ELEMENTS = 16
TEST_TILE_DOFS = 20
PAIR_TEST_TILE_ELEMENTS = 8  # single + double layer in one pass (BEM_TILE_PAIR_TEST_*)
PAIR_TEST_TILE_DOFS = 12
SYM_TEST_TILE_ELEMENTS = 8  # symmetric single + double layer pass (BEM_TILE_SYM_TEST_*)
SYM_TEST_TILE_DOFS = 12
SYM_RUN_ELEMENTS = 48  # test elements per work item (BEM_TILE_SYM_RUN_ELEMENTS)
TRIAL_BLOCK_ELEMENTS = 256
TRIAL_BLOCK_DOFS = 160
RWG_MIXED_TEST_TILE_ELEMENTS = 8  # mixed-precision EFIE (BEM_TILE_RWG_MIXED_*)
RWG_MIXED_TEST_TILE_DOFS = 16
RWG_MIXED_TRIAL_BLOCK_DOFS = 416
MAX_INCIDENCE = 12

_CACHE = {}


def _morton_keys(points):
    """63-bit Morton codes of points (n, 3) in their bounding box."""
    if len(points) == 0:
        return _np.zeros(0, dtype=_np.uint64)
    lo = points.min(axis=0)
    span = points.max(axis=0) - lo
    span[span == 0] = 1.0
    q = ((points - lo) / span * (2**21 - 1)).astype(_np.uint64)
    key = _np.zeros(len(points), dtype=_np.uint64)
    for bit in range(21):
        for d in range(3):
            key |= ((q[:, d] >> _np.uint64(bit)) & _np.uint64(1)) << _np.uint64(3 * bit + d)
    return key


class Blocks(object):
    """Host arrays of one blocking."""

    def __init__(self, order, start, dofs, ndofs, slots, max_dofs, nshape):
        self.order = order  # element indices in list order
        self.start = start
        self.dofs = dofs
        self.ndofs = ndofs
        self.slots = slots
        self.max_dofs = max_dofs
        self.nshape = nshape
        self.incidence = None

    def build_incidence(self, multipliers, max_incidence):
        """int16 [n_blocks, max_dofs, max_incidence]: codes 3 * (pos - start) + i of each slot, -1 padded."""
        inc = _np.full((self.n_blocks, self.max_dofs, max_incidence), -1, dtype=_np.int16)
        fill = _np.zeros((self.n_blocks, self.max_dofs), dtype=_np.int64)
        for b in range(self.n_blocks):
            for pos in range(self.start[b], self.start[b + 1]):
                e = self.order[pos]
                for i in range(self.nshape):
                    if multipliers[e, i] == 0:
                        continue
                    s = self.slots[pos, i]
                    k = fill[b, s]
                    if k >= max_incidence:
                        raise ValueError("a dof has more element incidences in one block than the device supports")
                    inc[b, s, k] = 3 * (pos - self.start[b]) + i
                    fill[b, s] = k + 1
        # keep only the incidence columns some dof uses (the kernel loops over them)
        used = max(1, int(fill.max()) if fill.size else 1)
        self.incidence = _np.ascontiguousarray(inc[:, :, :used])
        return self.incidence

    @property
    def n_blocks(self):
        return len(self.ndofs)

    @property
    def total_slots(self):
        return int(self.ndofs.sum())


def _greedy_blocks(order, l2g, mult, max_elements, max_dofs, max_incidence=MAX_INCIDENCE):
    nshape = l2g.shape[1]
    starts = [0]
    block_dofs = []
    slots = _np.zeros((len(order), 3), dtype=_np.int32)
    current = {}
    uses = {}
    count = 0
    for pos, e in enumerate(order):
        dofs = [int(l2g[e, i]) for i in range(nshape) if mult[e, i] != 0]
        new = len({d for d in dofs if d not in current})
        crowded = any(uses.get(d, 0) + dofs.count(d) > max_incidence for d in dofs)
        if count == max_elements or len(current) + new > max_dofs or crowded:
            block_dofs.append(current)
            starts.append(pos)
            current = {}
            uses = {}
            count = 0
        for i in range(nshape):
            if mult[e, i] != 0:
                d = int(l2g[e, i])
                if d not in current:
                    current[d] = len(current)
                uses[d] = uses.get(d, 0) + 1
                slots[pos, i] = current[d]
        count += 1
    block_dofs.append(current)
    starts.append(len(order))
    nb = len(block_dofs)
    dofs = _np.zeros((nb, max_dofs), dtype=_np.int32)
    ndofs = _np.zeros(nb, dtype=_np.int32)
    for b, cur in enumerate(block_dofs):
        # slots in ascending dof order: consecutive lanes of the flush then write
        # neighbouring columns of a matrix row, which share cache lines
        sorted_dofs = sorted(cur)
        rank = {d: s for s, d in enumerate(sorted_dofs)}
        ndofs[b] = len(cur)
        dofs[b, : len(cur)] = sorted_dofs
        old_to_new = _np.zeros(max(len(cur), 1), dtype=_np.int32)
        for d, s in cur.items():
            old_to_new[s] = rank[d]
        seg = slice(starts[b], starts[b + 1])
        slots[seg] = old_to_new[slots[seg]]
    return Blocks(_np.asarray(order, dtype=_np.int32), _np.asarray(starts, dtype=_np.int32), dofs, ndofs, slots,
                  max_dofs, nshape)


def make_blocks(space, max_elements, max_dofs):
    """Best of the grid order and the Morton order of centroids, by total slot count."""
    key = (space.id, max_elements, max_dofs)
    if key in _CACHE:
        return _CACHE[key]
    elements = space.support_elements.astype(_np.int64)
    l2g = space.local2global
    mult = space.local_multipliers
    natural = _greedy_blocks(elements, l2g, mult, max_elements, max_dofs)
    morton_order = elements[_np.argsort(_morton_keys(space.grid.centroids[elements]), kind="stable")]
    morton = _greedy_blocks(morton_order, l2g, mult, max_elements, max_dofs)
    best = natural if natural.total_slots <= morton.total_slots else morton
    _CACHE[key] = best
    return best


class DeviceBlocks(object):
    """A blocking in HBM with the space descriptor over its element order."""

    def __init__(self, space, blocks):
        import torch
        from bempp_amd.core import hip_backend as hb

        dev = hb.device()
        base = hb.device_space(space)
        self.blocks = blocks
        self.elements = torch.from_numpy(blocks.order).to(dev)
        self.start = torch.from_numpy(blocks.start).to(dev)
        self.dofs = torch.from_numpy(blocks.dofs).to(dev)
        self.ndofs = torch.from_numpy(blocks.ndofs).to(dev)
        self.slots = torch.from_numpy(blocks.slots).to(dev)
        if blocks.incidence is None:
            blocks.build_incidence(space.local_multipliers, MAX_INCIDENCE)
        self.incidence = torch.from_numpy(_np.ascontiguousarray(blocks.incidence)).to(dev)
        self.base = base
        self.space_struct = hb.BemSpace(
            int(len(blocks.order)), base.shapeset, self.elements.data_ptr(), base.grid.geometry.data_ptr(),
            base.grid.vertex_ids.data_ptr(), base.local2global.data_ptr(), base.multipliers.data_ptr(),
            base.normal_multipliers.data_ptr(),
        )
        self.struct = hb.BemDofBlocks(
            blocks.n_blocks, blocks.max_dofs, self.start.data_ptr(), self.dofs.data_ptr(), self.ndofs.data_ptr(),
            self.slots.data_ptr(), self.incidence.data_ptr(), int(blocks.incidence.shape[2]),
        )


_DEVICE_CACHE = {}


def device_blocks(space, role):
    """Device blocking of a space as test ('test', 'test_pair' for the fused pair, 'test_rwg_mixed')
    or trial ('trial', 'trial_rwg_mixed') side."""
    if role == "test":
        me, md = TEST_TILE_ELEMENTS, TEST_TILE_DOFS
    elif role == "test_pair":
        me, md = PAIR_TEST_TILE_ELEMENTS, PAIR_TEST_TILE_DOFS
    elif role == "test_rwg_mixed":
        me, md = RWG_MIXED_TEST_TILE_ELEMENTS, RWG_MIXED_TEST_TILE_DOFS
    elif role == "trial_rwg_mixed":
        me, md = TRIAL_BLOCK_ELEMENTS, RWG_MIXED_TRIAL_BLOCK_DOFS
    else:
        me, md = TRIAL_BLOCK_ELEMENTS, TRIAL_BLOCK_DOFS
    key = (space.id, me, md)
    if key not in _DEVICE_CACHE:
        _DEVICE_CACHE[key] = DeviceBlocks(space, make_blocks(space, me, md))
    return _DEVICE_CACHE[key]


def _tiles_per_item(n_pairs):
    """Test tiles per work item: runs as long as possible (the trial state is set up once per
    item) while leaving >= ~16k items, i.e. ~30 workgroups per resident slot on 256 CUs."""
    return int(max(1, min(32, n_pairs // 16384)))


def make_symmetric_blocks(space, tiles_per_item=None, trial_dofs=TRIAL_BLOCK_DOFS, min_tiles_per_item=1):
    """Blocking of the symmetric pass: (trial blocks, nested test tiles, tile -> block, work list).

    The trial blocks are those of make_blocks (TRIAL_BLOCK_ELEMENTS elements, at most trial_dofs
    dofs: TRIAL_BLOCK_DOFS, or RWG_MIXED_TRIAL_BLOCK_DOFS for the mixed electric field
    operator; min_tiles_per_item raises the tiles per work item above the default rule). Every
    trial block is cut into test tiles of at
    most SYM_TEST_TILE_ELEMENTS elements and SYM_TEST_TILE_DOFS dofs over the same element
    order, so each tile lies inside one block. Block b meets the tiles of blocks 0..b, a
    contiguous range; the work list (int32 [n, 3]: block, first tile, end tile) cuts that range
    into runs. With the kernel's rule for the tile's own block (only trial elements later in
    the order), each unordered element pair is visited once.
    """
    key = ("sym", space.id, tiles_per_item, trial_dofs, min_tiles_per_item)
    if key in _CACHE:
        return _CACHE[key]
    trial = make_blocks(space, TRIAL_BLOCK_ELEMENTS, trial_dofs)
    order = trial.order
    l2g = space.local2global
    mult = space.local_multipliers
    n = len(order)
    starts, dofs_rows, ndofs, owner = [], [], [], []
    slots = _np.zeros((n, 3), dtype=_np.int32)
    for b in range(trial.n_blocks):
        lo, hi = int(trial.start[b]), int(trial.start[b + 1])
        sub = _greedy_blocks(order[lo:hi], l2g, mult, SYM_TEST_TILE_ELEMENTS, SYM_TEST_TILE_DOFS)
        starts.extend(int(lo + st) for st in sub.start[:-1])
        dofs_rows.append(sub.dofs)
        ndofs.append(sub.ndofs)
        slots[lo:hi] = sub.slots
        owner.extend([b] * sub.n_blocks)
    starts.append(n)
    tiles = Blocks(order.copy(), _np.asarray(starts, dtype=_np.int32),
                   _np.concatenate(dofs_rows) if dofs_rows else _np.zeros((0, SYM_TEST_TILE_DOFS), _np.int32),
                   _np.concatenate(ndofs) if ndofs else _np.zeros(0, _np.int32), slots, SYM_TEST_TILE_DOFS,
                   trial.nshape)
    owner = _np.asarray(owner, dtype=_np.int64)
    # tiles of blocks 0..b: [0, tiles_end[b])
    tiles_end = _np.searchsorted(owner, _np.arange(trial.n_blocks), side="right")
    g = tiles_per_item or max(min_tiles_per_item, _tiles_per_item(int(tiles_end.sum())))
    g = min(g, SYM_RUN_ELEMENTS // SYM_TEST_TILE_ELEMENTS)
    items = [(b, t, min(t + g, int(te))) for b, te in enumerate(tiles_end) for t in range(0, int(te), g)]
    items.sort(key=lambda it: (it[1], it[0]))  # concurrent workgroups share test tiles (output rows)
    work = _np.asarray(items, dtype=_np.int32).reshape(-1, 3)
    result = (trial, tiles, owner, work)
    _CACHE[key] = result
    return result


class DeviceSymmetricBlocks(object):
    """Device arrays of make_symmetric_blocks: trial blocks (DeviceBlocks), test tiles, work list.
    trial_role: 'trial', or 'trial_rwg_mixed' for the mixed electric field operator."""

    def __init__(self, space, trial_role="trial"):
        import torch
        from bempp_amd.core import hip_backend as hb

        dev = hb.device()
        # the mixed electric field pass stages ~400 edges of trial data per work item: at least
        # two test tiles per item amortise it (one tile per item is ~20% slower at 5120 triangles)
        rwgm = trial_role == "trial_rwg_mixed"
        _, tiles, owner, work = make_symmetric_blocks(
            space, trial_dofs=RWG_MIXED_TRIAL_BLOCK_DOFS if rwgm else TRIAL_BLOCK_DOFS,
            min_tiles_per_item=2 if rwgm else 1)
        self.trial = device_blocks(space, trial_role)
        self.tiles = tiles
        self.tile_block = owner
        self.start = torch.from_numpy(tiles.start).to(dev)
        self.dofs = torch.from_numpy(_np.ascontiguousarray(tiles.dofs)).to(dev)
        self.ndofs = torch.from_numpy(tiles.ndofs).to(dev)
        self.slots = torch.from_numpy(tiles.slots).to(dev)
        self.work = torch.from_numpy(work).to(dev)
        self.n_work = int(work.shape[0])
        self.tiles_struct = hb.BemDofBlocks(
            tiles.n_blocks, tiles.max_dofs, self.start.data_ptr(), self.dofs.data_ptr(), self.ndofs.data_ptr(),
            self.slots.data_ptr(), None, 0,
        )


def device_symmetric_blocks(space, trial_role="trial"):
    key = ("sym", space.id, trial_role)
    if key not in _DEVICE_CACHE:
        _DEVICE_CACHE[key] = DeviceSymmetricBlocks(space, trial_role)
    return _DEVICE_CACHE[key]


def clear_cache():
    _CACHE.clear()
    _DEVICE_CACHE.clear()
