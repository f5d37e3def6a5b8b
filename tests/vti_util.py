"""Test helpers for the VTI format of the PDFS CLI (single_gpu_cuda/src/
vti_parser.c input layout; file_writer.c output layout)."""
import numpy as np


def write_input_vti(path, loc, dist, h=(1.0, 1.0, 1.0)):
    """Input file in the layout the reference parser reads: 6 header lines
    (the 3rd holds WholeExtent then Spacing), interior location ints, 2
    lines, interior distance values, closing lines. loc/dist: [Z][Y][X]."""
    Z, Y, X = loc.shape
    lines = ['<?xml version="1.0"?>',
             '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">',
             f'    <ImageData WholeExtent="0 {X - 1} 0 {Y - 1} 0 {Z - 1}" '
             f'Spacing="{h[0]!r} {h[1]!r} {h[2]!r}">',
             f'  <Piece Extent="0 {X - 1} 0 {Y - 1} 0 {Z - 1}">',
             '    <PointData Scalars="location">',
             '      <DataArray type="Int32" Name="location" format="ascii">']
    body = [" ".join(str(int(v)) for v in row) for row in loc.reshape(-1, X)]
    mid = ['      </DataArray>',
           '      <DataArray type="Float64" Name="distance" format="ascii">']
    dbody = [" ".join(repr(float(v)) for v in row) for row in dist.reshape(-1, X)]
    tail = ['      </DataArray>', '    </PointData>', '  </Piece>', '  </ImageData>',
            '</VTKFile>']
    with open(path, "w") as f:
        f.write("\n".join(lines + body + mid + dbody + tail) + "\n")


def reference_vti_text(padded, h):
    """file_writer.c:152-190 restated in Python (the expected CLI output)."""
    gz, gy, gx = padded.shape
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">\n',
           '\t<ImageData WholeExtent="0 %d 0 %d 0 %d" Origin="0 0 0" Spacing=" %f %f %f  ">\n'
           % (gx - 1, gy - 1, gz - 1, h[0], h[1], h[2]),
           '\t\t<Piece Extent="0 %d 0 %d 0 %d">\n' % (gx - 1, gy - 1, gz - 1),
           '\t\t\t<PointData Scalars="Distance Field">\n',
           '\t\t\t\t<DataArray type="Float32" Name="Distance Field" format="ascii">\n']
    for i in range(gz):
        for j in range(gy):
            out.append("".join("%f  " % v for v in padded[i, j]) + "\n")
        out.append("\n")
    out += ['\t\t\t\t</DataArray>\n', '\t\t\t</PointData>\n', '\t\t\t<CellData>\n',
            '\t\t\t</CellData>\n', '\t\t</Piece>\n', '\t</ImageData>\n', '</VTKFile>\n']
    return "".join(out)


def random_cli_input(X, Y, Z, seed):
    """Location flags (0 outside, 1 inside) and distances with exact zeros
    (unknown), NaN, negatives and positives, as CLI inputs hold them."""
    rng = np.random.default_rng(seed)
    loc = (rng.uniform(size=(Z, Y, X)) < 0.1).astype(np.int32)
    dist = rng.uniform(0.0, 5.0, size=(Z, Y, X))
    dist[rng.uniform(size=(Z, Y, X)) < 0.6] = 0.0
    dist[loc == 1] = rng.uniform(-2.0, 0.0, size=int((loc == 1).sum()))
    return loc, dist
