"""Surface grids of flat triangles: geometry and adjacency.

Counterpart of bempp_cl/api/grid/grid.py:15-634. A Grid keeps
* vertices (3, V) float64 and elements (3, F) uint32 (Fortran order, as the reference);
* edges (2, E) and element_edges (3, F), enumerated in order of first occurrence over
  elements and their local edges (0,1), (2,0), (1,2) (grid.py:12, 1408-1440), which fixes
  the RWG/SNC dof numbering;
* per-element geometry (grid.py:552-590): jacobians (F,3,2) = [v1-v0, v2-v0], unit
  normals, integration elements |(v1-v0) x (v2-v0)|, volumes, centroids, diameters;
* edge adjacency (6, n) and vertex adjacency (4, n) of ordered element pairs with the
  local vertex indices of the shared edge/vertex (grid.py:116-140, 1015-1098).

Everything is computed with vectorised numpy; nothing here runs on the GPU.
"""

import itertools as _itertools

import numpy as _np

_EDGE_LOCAL = _np.array([[0, 1], [2, 0], [1, 2]])

_ID_COUNTER = _itertools.count()


class GridData(object):
    """Plain container of the arrays kernels need (grid.py:636-697)."""

    def __init__(self, grid, dtype):
        self.vertices = grid.vertices.astype(dtype)
        self.elements = grid.elements
        self.normals = grid.normals.astype(dtype)
        self.jacobians = grid.jacobians.astype(dtype)
        self.integration_elements = grid.integration_elements.astype(dtype)
        self.volumes = grid.volumes.astype(dtype)
        self.centroids = grid.centroids.astype(dtype)
        self.diameters = grid.diameters.astype(dtype)
        self.vertex_on_boundary = grid.vertex_on_boundary

    def local2global(self, elem_index, local_coords):
        """Map reference coordinates (2, N) of one element to global points (3, N)."""
        return self.vertices[:, self.elements[0, elem_index]][:, None] + self.jacobians[elem_index] @ local_coords


class Grid(object):
    """A surface grid of flat triangles."""

    def __init__(self, vertices, elements, domain_indices=None, grid_id=None):
        """Create a grid from a (3, V) vertex array and a (3, F) element array."""
        self._vertices = _np.asfortranarray(_np.asarray(vertices, dtype=_np.float64))
        self._elements = _np.asfortranarray(_np.asarray(elements, dtype=_np.uint32))
        if self._vertices.ndim != 2 or self._vertices.shape[0] != 3:
            raise ValueError("vertices must have shape (3, n_vertices)")
        if self._elements.ndim != 2 or self._elements.shape[0] != 3:
            raise ValueError("elements must have shape (3, n_elements)")
        if domain_indices is None:
            domain_indices = _np.zeros(self._elements.shape[1], dtype=_np.uint32)
        self._domain_indices = _np.asarray(domain_indices, dtype=_np.uint32)
        self._id = grid_id if grid_id else f"grid{next(_ID_COUNTER)}"
        self._topology_id = self._id
        self._data = {}

        self._enumerate_edges()
        self._compute_geometry()
        self._compute_adjacency()
        self._compute_boundary_information()

    def with_vertices(self, vertices):
        """A new grid with the same elements and moved vertices.

        Edges, adjacency and boundary information depend on the topology only and are
        shared; the geometry (jacobians, normals, diameters, ...) is recomputed.
        """
        v = _np.asfortranarray(_np.asarray(vertices, dtype=_np.float64))
        if v.shape != self._vertices.shape:
            raise ValueError(f"vertices must have shape {self._vertices.shape}")
        g = object.__new__(Grid)
        g.__dict__.update(self.__dict__)
        g._vertices = v
        g._id = f"grid{next(_ID_COUNTER)}"
        g._data = {}
        g._compute_geometry()
        return g

    # ------------------------------------------------------------------ setup
    def _enumerate_edges(self):
        """Number edges by first occurrence (grid.py:1408-1440)."""
        nel = self.number_of_elements
        # local edge l of element e is elements[_EDGE_LOCAL[l], e], sorted ascending
        pairs = self._elements[_EDGE_LOCAL.T]  # (2, 3, F)
        lo = _np.minimum(pairs[0], pairs[1]).T.ravel().astype(_np.int64)  # order (e, l)
        hi = _np.maximum(pairs[0], pairs[1]).T.ravel().astype(_np.int64)
        keys = lo * (int(self._vertices.shape[1]) + 1) + hi
        uniq, first, inverse = _np.unique(keys, return_index=True, return_inverse=True)
        # rank unique edges by their first occurrence
        order = _np.argsort(first, kind="stable")
        rank = _np.empty_like(order)
        rank[order] = _np.arange(len(order))
        edge_index = rank[inverse.ravel()]
        self._edges = _np.empty((2, len(uniq)), dtype=_np.uint32)
        self._edges[0, edge_index] = lo
        self._edges[1, edge_index] = hi
        self._element_edges = _np.asfortranarray(edge_index.reshape(nel, 3).T.astype(_np.uint32))

    def _compute_geometry(self):
        """Jacobians, normals, integration elements and derived data (grid.py:552-590)."""
        v = self._vertices
        el = self._elements
        v0 = v[:, el[0]].T
        v1 = v[:, el[1]].T
        v2 = v[:, el[2]].T
        j0 = v1 - v0
        j1 = v2 - v0
        cross = _np.cross(j0, j1)
        norms = _np.linalg.norm(cross, axis=1)
        if _np.any(norms == 0):
            raise ValueError("Grid contains degenerate elements.")
        self._normals = cross / norms[:, None]
        self._volumes = 0.5 * norms
        self._centroids = (v0 + v1 + v2) / 3.0
        self._jacobians = _np.stack([j0, j1], axis=2)  # (F, 3, 2)
        self._integration_elements = norms
        diff = _np.linalg.norm(j0 - j1, axis=1)
        self._diameters = _np.linalg.norm(j0, axis=1) * _np.linalg.norm(j1, axis=1) * diff / norms
        jtj = _np.einsum("fki,fkj->fij", self._jacobians, self._jacobians)
        self._jacobian_inverse_transposed = _np.einsum("fki,fij->fkj", self._jacobians, _np.linalg.inv(jtj))

    def _compute_adjacency(self):
        """Ordered element pairs sharing an edge or exactly one vertex (grid.py:510-550)."""
        from scipy.sparse import csr_matrix

        nel = self.number_of_elements
        nv = self.number_of_vertices
        el = self._elements
        rows = el.T.ravel().astype(_np.int64)
        cols = _np.repeat(_np.arange(nel), 3)
        vert_to_elem = csr_matrix((_np.ones(3 * nel, dtype=_np.int32), (rows, cols)), shape=(nv, nel))
        self._element_to_vertex_matrix = vert_to_elem
        e2e = (vert_to_elem.T @ vert_to_elem).tocoo()
        self._element_to_element_matrix = e2e.tocsr()
        off = e2e.row != e2e.col
        r = e2e.row[off]
        c = e2e.col[off]
        cnt = e2e.data[off]
        order = _np.lexsort((c, r))
        r, c, cnt = r[order], c[order], cnt[order]

        # local positions of shared vertices: match[p, a, b] = el[a, r] == el[b, c]
        match = el[:, r][:, None, :] == el[:, c][None, :, :]  # (3, 3, n)
        match = _np.moveaxis(match, 2, 0)

        vsel = cnt == 1
        mv = match[vsel]
        flat = mv.reshape(len(mv), 9).argmax(axis=1)
        self._vertex_adjacency = _np.vstack([r[vsel], c[vsel], flat // 3, flat % 3]).astype(_np.int32)

        esel = cnt == 2
        me = match[esel]
        # test-side positions in increasing order; their trial-side partners
        in_test = me.any(axis=2)  # (n, 3)
        i_sorted = _np.argsort(~in_test, axis=1, kind="stable")[:, :2]
        ne = len(me)
        idx = _np.arange(ne)
        j0 = me[idx, i_sorted[:, 0]].argmax(axis=1)
        j1 = me[idx, i_sorted[:, 1]].argmax(axis=1)
        i0 = i_sorted[:, 0].copy()
        i1 = i_sorted[:, 1].copy()
        swap = j1 < j0  # grid.py:1036-1044
        i0[swap], i1[swap] = i_sorted[swap, 1], i_sorted[swap, 0]
        j0s = _np.where(swap, j1, j0)
        j1s = _np.where(swap, j0, j1)
        self._edge_adjacency = _np.vstack([r[esel], c[esel], i0, i1, j0s, j1s]).astype(_np.int32)

        # vertex -> elements lists (CSR over vertices)
        vt = vert_to_elem.tocsr()
        vt.sort_indices()
        self._vertex_neighbors = (vt.indices.astype(_np.uint32), vt.indptr.astype(_np.uint32))

        # edge -> element lists
        nedges = self._edges.shape[1]
        ee = self._element_edges.T.ravel()
        owner = _np.repeat(_np.arange(nel), 3)
        order = _np.argsort(ee, kind="stable")
        counts = _np.bincount(ee, minlength=nedges)
        ptr = _np.concatenate([[0], _np.cumsum(counts)])
        self._edge_neighbors_flat = (owner[order].astype(_np.uint32), ptr.astype(_np.uint32))

    def _compute_boundary_information(self):
        """Edges with one neighbour are boundary edges; their vertices are boundary vertices."""
        nedges = self._edges.shape[1]
        counts = _np.bincount(self._element_edges.ravel(), minlength=nedges)
        self._edge_on_boundary = counts == 1
        self._vertex_on_boundary = _np.zeros(self.number_of_vertices, dtype=bool)
        self._vertex_on_boundary[self._edges[:, self._edge_on_boundary].ravel()] = True

    # ------------------------------------------------------------- properties
    @property
    def topology_id(self):
        """Identifier shared by grids with the same elements (see with_vertices)."""
        return self._topology_id

    @property
    def id(self):
        return self._id

    @property
    def vertices(self):
        return self._vertices

    @property
    def elements(self):
        return self._elements

    @property
    def domain_indices(self):
        return self._domain_indices

    @property
    def edges(self):
        return self._edges

    @property
    def element_edges(self):
        return self._element_edges

    @property
    def number_of_vertices(self):
        return self._vertices.shape[1]

    @property
    def number_of_elements(self):
        return self._elements.shape[1]

    @property
    def number_of_edges(self):
        return self._edges.shape[1]

    @property
    def normals(self):
        return self._normals

    @property
    def jacobians(self):
        return self._jacobians

    @property
    def jacobian_inverse_transposed(self):
        return self._jacobian_inverse_transposed

    @property
    def integration_elements(self):
        return self._integration_elements

    @property
    def volumes(self):
        return self._volumes

    @property
    def centroids(self):
        return self._centroids

    @property
    def diameters(self):
        return self._diameters

    @property
    def edge_adjacency(self):
        """(6, n): e0, e1, local shared vertices in e0 (2), matching local vertices in e1 (2)."""
        return self._edge_adjacency

    @property
    def vertex_adjacency(self):
        """(4, n): e0, e1, local index of the shared vertex in e0 and in e1."""
        return self._vertex_adjacency

    @property
    def element_to_vertex_matrix(self):
        return self._element_to_vertex_matrix

    @property
    def element_to_element_matrix(self):
        return self._element_to_element_matrix

    @property
    def vertex_neighbors(self):
        """(indices, indexptr) of the elements around each vertex."""
        return self._vertex_neighbors

    @property
    def edge_neighbors(self):
        """Tuple of element tuples adjacent to each edge (grid.py:610-616)."""
        idx, ptr = self._edge_neighbors_flat
        return [tuple(int(x) for x in idx[ptr[e] : ptr[e + 1]]) for e in range(self.number_of_edges)]

    @property
    def edge_neighbors_flat(self):
        return self._edge_neighbors_flat

    @property
    def vertex_on_boundary(self):
        return self._vertex_on_boundary

    @property
    def edge_on_boundary(self):
        return self._edge_on_boundary

    def data(self, precision="double"):
        """Return a GridData container in the given precision."""
        if precision not in ("single", "double"):
            raise ValueError("precision must be one of 'single' or 'double'")
        if precision not in self._data:
            self._data[precision] = GridData(self, _np.float64 if precision == "double" else _np.float32)
        return self._data[precision]

    def as_array(self):
        """Corner coordinates, 9 values per element: v0, v1, v2 (grid.py:227-236)."""
        return self._vertices.T[self._elements.ravel(order="F")].ravel()

    def refine(self):
        """Split every triangle into 4 through its edge midpoints (grid.py:431-464).

        Child order per element: (v0, m01, m20), (m01, v1, m12), (m12, v2, m20), (m01, m12, m20),
        where midpoint vertices are numbered V + edge index.
        """
        nv = self.number_of_vertices
        new_vertices = _np.hstack([self._vertices, 0.5 * (self._vertices[:, self._edges[0]] + self._vertices[:, self._edges[1]])])
        el = self._elements.astype(_np.int64)
        m01 = self._element_edges[0].astype(_np.int64) + nv
        m20 = self._element_edges[1].astype(_np.int64) + nv
        m12 = self._element_edges[2].astype(_np.int64) + nv
        children = _np.stack(
            [
                _np.stack([el[0], m01, m20]),
                _np.stack([m01, el[1], m12]),
                _np.stack([m12, el[2], m20]),
                _np.stack([m01, m12, m20]),
            ],
            axis=2,
        )  # (3, F, 4)
        new_elements = children.reshape(3, -1)
        return Grid(new_vertices, new_elements, _np.repeat(self._domain_indices, 4))

    def __eq__(self, other):
        return isinstance(other, Grid) and self._id == other._id

    def __hash__(self):
        return hash(self._id)
