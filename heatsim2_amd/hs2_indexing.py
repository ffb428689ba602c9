"""Index maps for compressed geometry: z-layers whose x-y resolution is
reduced by a power of two per layer.

The upstream module (hs2_indexing.py:9-63) is marked unused and none of its
functions runs as shipped (it takes len() of a function, reads an undefined
name, and two functions return nothing).  This module implements the
evident intent: layer k keeps (ny >> s[k]) x (nx >> s[k]) cells, a
compressed cell (k, j', i') covers uncompressed cells
(k, j' << s[k] .. ((j'+1) << s[k]) - 1, i' likewise), and compressed cells
are numbered layer by layer in C order.
"""


def _layer_shape(xy_scaling_log2, uncompressed_shape, k):
    s = xy_scaling_log2[k]
    return uncompressed_shape[1] >> s, uncompressed_shape[2] >> s


def compressed_index(uncompressed_index, xy_scaling_log2):
    """(k, j, i) -> (k, j >> s[k], i >> s[k])."""
    if len(uncompressed_index) != 3:
        raise ValueError("need a (k, j, i) index")
    k, j, i = uncompressed_index
    s = xy_scaling_log2[k]
    return (k, j >> s, i >> s)


def compressed_single_index(compressed_idx, xy_scaling_log2, uncompressed_shape):
    """Flat position of compressed cell (k, j', i')."""
    k, j, i = compressed_idx
    off = 0
    for z in range(k):
        cy, cx = _layer_shape(xy_scaling_log2, uncompressed_shape, z)
        off += cy * cx
    cy, cx = _layer_shape(xy_scaling_log2, uncompressed_shape, k)
    if not (0 <= j < cy and 0 <= i < cx):
        raise IndexError("compressed index %r outside layer %d (%d x %d)" % (compressed_idx, k, cy, cx))
    return off + j * cx + i


def compressed_index_from_single(single_index, xy_scaling_log2, uncompressed_shape):
    """Inverse of compressed_single_index."""
    off = 0
    for k in range(uncompressed_shape[0]):
        cy, cx = _layer_shape(xy_scaling_log2, uncompressed_shape, k)
        if single_index < off + cy * cx:
            rem = single_index - off
            return (k, rem // cx, rem % cx)
        off += cy * cx
    raise IndexError("flat index %d beyond the compressed grid (%d cells)" % (single_index, off))


def uncompressed_index_range(compressed_idx, xy_scaling_log2):
    """((k, k+1), (j0, j1), (i0, i1)) half-open ranges covered by a
    compressed cell."""
    k, j, i = compressed_idx
    f = 1 << xy_scaling_log2[k]
    return ((k, k + 1), (j * f, (j + 1) * f), (i * f, (i + 1) * f))
