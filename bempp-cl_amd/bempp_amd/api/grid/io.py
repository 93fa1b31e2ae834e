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


def export(filename, grid=None):
    """Write a grid as an ASCII gmsh MSH 2.2 file (triangles with physical tag = domain index)."""
    if grid is None:
        raise ValueError("A grid must be given.")
    with open(filename, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n")
        f.write(f"{grid.number_of_vertices}\n")
        for i, v in enumerate(grid.vertices.T):
            f.write(f"{i + 1} {float(v[0])!r} {float(v[1])!r} {float(v[2])!r}\n")
        f.write("$EndNodes\n$Elements\n")
        f.write(f"{grid.number_of_elements}\n")
        for i, (e, d) in enumerate(zip(grid.elements.T, grid.domain_indices)):
            f.write(f"{i + 1} 2 2 {d} {d} {e[0] + 1} {e[1] + 1} {e[2] + 1}\n")
        f.write("$EndElements\n")
