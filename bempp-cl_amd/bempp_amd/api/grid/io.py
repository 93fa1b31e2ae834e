"""Grid import: gmsh MSH 2.2 files (ASCII and binary).

Counterpart of bempp_cl/api/grid/io.py:7-36, which delegates to meshio. meshio is not
part of this library; this module reads the MSH 2.2 format directly and applies the
same conventions: triangles only, node tags mapped to consecutive vertex indices in
file order, and domain indices taken from the physical tag, or from the geometrical tag
when every physical tag is zero.
"""

import struct as _struct

import numpy as _np

# nodes per element for the gmsh element types that can appear in surface meshes
_NODES_PER_TYPE = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5, 8: 3, 9: 6, 10: 9, 11: 10, 15: 1}


def _read_msh22(data):
    pos = 0

    def readline():
        nonlocal pos
        end = data.index(b"\n", pos)
        line = data[pos:end].strip()
        pos = end + 1
        return line

    header = None
    binary = False
    endian = "<"
    node_tags = None
    coords = None
    triangles = []
    tags_phys = []
    tags_geom = []
    while pos < len(data):
        line = readline()
        if line == b"$MeshFormat":
            fields = readline().split()
            header = fields
            if fields[0] not in (b"2.2", b"2.1", b"2"):
                raise ValueError("Only gmsh MSH 2.x files are supported.")
            binary = fields[1] == b"1"
            if binary:
                one = data[pos : pos + 4]
                endian = "<" if _struct.unpack("<i", one)[0] == 1 else ">"
                pos += 4
                readline()
            readline()  # $EndMeshFormat
        elif line == b"$Nodes":
            n = int(readline())
            if binary:
                rec = _np.dtype([("tag", endian + "i4"), ("x", endian + "f8", (3,))])
                arr = _np.frombuffer(data, dtype=rec, count=n, offset=pos)
                pos += n * rec.itemsize
                node_tags = arr["tag"].astype(_np.int64)
                coords = arr["x"].astype(_np.float64)
                readline()
            else:
                vals = _np.array([readline().split() for _ in range(n)], dtype=_np.float64)
                node_tags = vals[:, 0].astype(_np.int64)
                coords = vals[:, 1:4]
            readline()  # $EndNodes
        elif line == b"$Elements":
            n = int(readline())
            if binary:
                done = 0
                while done < n:
                    etype, nfollow, ntags = _struct.unpack(endian + "3i", data[pos : pos + 12])
                    pos += 12
                    nn = _NODES_PER_TYPE[etype]
                    width = 1 + ntags + nn
                    block = _np.frombuffer(data, dtype=endian + "i4", count=nfollow * width, offset=pos)
                    pos += 4 * nfollow * width
                    block = block.reshape(nfollow, width)
                    if etype == 2:
                        triangles.append(block[:, 1 + ntags :])
                        tags_phys.append(block[:, 1] if ntags > 0 else _np.zeros(nfollow, _np.int64))
                        tags_geom.append(block[:, 2] if ntags > 1 else _np.zeros(nfollow, _np.int64))
                    done += nfollow
                readline()
            else:
                for _ in range(n):
                    f = [int(x) for x in readline().split()]
                    etype, ntags = f[1], f[2]
                    if etype == 2:
                        triangles.append(_np.array([f[3 + ntags : 6 + ntags]]))
                        tags_phys.append(_np.array([f[3] if ntags > 0 else 0]))
                        tags_geom.append(_np.array([f[4] if ntags > 1 else 0]))
            readline()  # $EndElements
        elif line.startswith(b"$") and not line.startswith(b"$End"):
            # skip unknown sections
            name = line[1:]
            while readline() != b"$End" + name:
                pass
    if header is None or coords is None:
        raise ValueError("Not a valid gmsh MSH 2.2 file.")
    tri = _np.concatenate(triangles).astype(_np.int64)
    lookup = _np.full(int(node_tags.max()) + 1, -1, dtype=_np.int64)
    lookup[node_tags] = _np.arange(len(node_tags))
    elements = lookup[tri].T
    if _np.any(elements < 0):
        raise ValueError("Element references an undefined node.")
    phys = _np.concatenate(tags_phys).astype(_np.uint32)
    geom = _np.concatenate(tags_geom).astype(_np.uint32)
    domain_indices = geom if _np.all(phys == 0) else phys
    return coords.T.copy(), elements, domain_indices


def import_grid(filename):
    """Read a gmsh MSH 2.2 file and return a Grid."""
    from bempp_amd.api.grid.grid import Grid

    with open(filename, "rb") as f:
        data = f.read()
    vertices, elements, domain_indices = _read_msh22(data)
    return Grid(vertices, elements, domain_indices=domain_indices)


def _transform_array(a, mode=None):
    """Transform data before export: None, 'real', 'imag', 'abs', 'abs_squared', 'log_abs' (the
    last three over the components, axis 0) or a callable on the (components, n) array
    (bempp_cl/api/grid/io.py:163-200)."""
    if mode is None:
        return a
    if mode == "real":
        return _np.real(a)
    if mode == "imag":
        return _np.imag(a)
    if mode == "abs":
        return _np.sqrt(_np.sum(_np.abs(a) ** 2, axis=0, keepdims=True))
    if mode == "abs_squared":
        return _np.sum(_np.abs(a) ** 2, axis=0, keepdims=True)
    if mode == "log_abs":
        return _np.log(_np.sqrt(_np.sum(_np.abs(a) ** 2, axis=0, keepdims=True)))
    if callable(mode):
        return mode(a)
    raise ValueError(f"Unknown transformation {mode!r}")


def _write_data_section(f, section, name, values, binary):
    """$NodeData / $ElementData of the MSH 2.2 format: one string tag (the view name), one real
    tag (time 0), three integer tags (time step 0, components, entries), then per entity its
    1-based tag and its components."""
    values = _np.asarray(values, dtype=_np.float64)
    if values.ndim == 1:
        values = values[:, None]
    n, ncomp = values.shape
    if ncomp not in (1, 3, 9):
        raise ValueError("MSH 2.2 data fields have 1, 3 or 9 components")
    f.write(f'${section}\n1\n"{name}"\n1\n0.0\n3\n0\n{ncomp}\n{n}\n'.encode())
    tags = _np.arange(1, n + 1, dtype="<i4")
    if binary:
        rec = _np.empty(n, dtype=[("tag", "<i4"), ("v", "<f8", (ncomp,))])
        rec["tag"] = tags
        rec["v"] = values
        f.write(rec.tobytes())
        f.write(b"\n")
    else:
        for t, row in zip(tags, values):
            f.write((f"{t} " + " ".join(repr(float(x)) for x in row) + "\n").encode())
    f.write(f"$End{section}\n".encode())


def export(filename, grid=None, grid_function=None, data_type=None, transformation=None, write_binary=True):
    """Export a grid or a grid function as a gmsh MSH 2.2 file (bempp_cl/api/grid/io.py:39-160).

    Exactly one of `grid` and `grid_function` is given. The file holds the vertices, the
    triangles with the domain index as physical tag and consecutive geometrical tags per domain
    index; for a grid function also its values: data_type 'node' (vertex values) or 'element'
    (values at the element centres), complex values as two fields 'real' and 'imag', otherwise
    one field 'data'. As in the reference, the default data_type is 'element' unless the space's
    identifier is 'p1' (its P1 space is called 'p1_continuous', so element data is the default).
    write_binary: binary MSH 2.2 (the default, as the reference) or ASCII. Only the .msh format
    is written.
    """
    import os

    if (grid is None) == (grid_function is None):
        raise ValueError("Exactly one of 'grid' and 'grid_function' must be supplied.")
    if os.path.splitext(filename)[1] != ".msh":
        raise ValueError("Only gmsh .msh (MSH 2.2) files are written by this library.")
    if data_type is None and grid_function is not None:
        data_type = "node" if grid_function.space.identifier == "p1" else "element"
    fields = []
    if grid_function is not None:
        grid = grid_function.space.grid
        if data_type == "node":
            data = _transform_array(grid_function.evaluate_on_vertices(), transformation).T
            section = "NodeData"
        elif data_type == "element":
            data = _transform_array(grid_function.evaluate_on_element_centers(), transformation).T
            section = "ElementData"
        else:
            raise ValueError("'data_type' must be one of 'element' or 'node'")
        if _np.iscomplexobj(data):
            fields = [(section, "real", _np.real(data)), (section, "imag", _np.imag(data))]
        else:
            fields = [(section, "data", data)]
    nv, ne = grid.number_of_vertices, grid.number_of_elements
    phys = _np.asarray(grid.domain_indices).astype(_np.int64)
    geom_of = {d: i + 1 for i, d in enumerate(sorted(set(phys.tolist())))}
    geom = _np.array([geom_of[d] for d in phys.tolist()], dtype=_np.int64)
    with open(filename, "wb") as f:
        f.write(b"$MeshFormat\n")
        if write_binary:
            f.write(b"2.2 1 8\n")
            f.write(_np.array([1], dtype="<i4").tobytes())
            f.write(b"\n")
        else:
            f.write(b"2.2 0 8\n")
        f.write(b"$EndMeshFormat\n$Nodes\n")
        f.write(f"{nv}\n".encode())
        if write_binary:
            rec = _np.empty(nv, dtype=[("tag", "<i4"), ("x", "<f8", (3,))])
            rec["tag"] = _np.arange(1, nv + 1)
            rec["x"] = grid.vertices.T
            f.write(rec.tobytes())
            f.write(b"\n")
        else:
            for i, v in enumerate(grid.vertices.T):
                f.write(f"{i + 1} {float(v[0])!r} {float(v[1])!r} {float(v[2])!r}\n".encode())
        f.write(b"$EndNodes\n$Elements\n")
        f.write(f"{ne}\n".encode())
        if write_binary:
            f.write(_np.array([2, ne, 2], dtype="<i4").tobytes())  # triangles, count, two tags
            body = _np.empty((ne, 6), dtype="<i4")
            body[:, 0] = _np.arange(1, ne + 1)
            body[:, 1] = phys
            body[:, 2] = geom
            body[:, 3:] = grid.elements.T + 1
            f.write(body.tobytes())
            f.write(b"\n")
        else:
            for i in range(ne):
                e = grid.elements[:, i]
                f.write(f"{i + 1} 2 2 {phys[i]} {geom[i]} {e[0] + 1} {e[1] + 1} {e[2] + 1}\n".encode())
        f.write(b"$EndElements\n")
        for section, name, values in fields:
            _write_data_section(f, section, name, values, write_binary)


def read_data_fields(filename):
    """The $NodeData / $ElementData fields of a MSH 2.2 file: {(section, name): (n, components) array}."""
    with open(filename, "rb") as f:
        data = f.read()
    binary = b"\n2.2 1 8\n" in data[:32]
    out = {}
    for section in (b"NodeData", b"ElementData"):
        start = 0
        while True:
            k = data.find(b"$" + section + b"\n", start)
            if k < 0:
                break
            pos = k + len(section) + 2
            lines = []
            while len(lines) < 8:  # string tags (1, name), real tags (1, time), integer tags (3, step, comps, n)
                end = data.index(b"\n", pos)
                lines.append(data[pos:end].strip())
                pos = end + 1
            name = lines[1].decode().strip('"')
            ncomp, n = int(lines[6]), int(lines[7])
            if binary:
                rec = _np.dtype([("tag", "<i4"), ("v", "<f8", (ncomp,))])
                arr = _np.frombuffer(data, dtype=rec, count=n, offset=pos)
                vals = _np.array(arr["v"]).reshape(n, ncomp)
                pos += n * rec.itemsize + 1
            else:
                rows = []
                for _ in range(n):
                    end = data.index(b"\n", pos)
                    rows.append([float(x) for x in data[pos:end].split()[1:]])
                    pos = end + 1
                vals = _np.array(rows).reshape(n, ncomp)
            out[(section.decode(), name)] = vals
            start = pos
    return out
