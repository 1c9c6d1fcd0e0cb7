"""File formats on either side of the feature path (host code of the
library, no GPU):
  * .sig, .fit and weights files that the reference's own tests hold are
    read back to the values of the golden fixtures and written again byte
    for byte (signatures.cpp:227-304, TrainingSet.cpp:339-480, 594-642);
    test-l.fit, written by an older "%.5e" writer, reads to its values;
  * the reference's TIFF inputs (uncompressed big-endian 16-bit, LZW 16-bit,
    8-bit strips) decode to the pixels of the golden fixtures, and TIFFs
    written by PIL (LZW / PackBits / Deflate, gray and RGB, predictor 2,
    several strips) to the arrays written; what LoadTIFF refuses (other
    sample sizes, several directories) is refused."""
import gzip
import os
import struct
import zlib

import numpy as np
import pytest

from helpers import GOLDEN, ROOT, load_npz

LIB = os.path.join(ROOT, 'wnd-charm_amd', 'lib', 'libwndgpu.so')
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason='library not built')
RAW = os.path.join(GOLDEN, 'raw')
LYMPHOMA = 'lymphoma_eosin_channel_MCL_test_img_sj-05-3362-R2_001_E-t6x5_5_4-l.tiff'


def raw(name, tmp_path):
    """the fixture's original bytes, and a path holding them"""
    data = gzip.open(os.path.join(RAW, name + '.gz'), 'rb').read()
    p = tmp_path / name
    p.write_bytes(data)
    return data, str(p)


@pytest.mark.parametrize('name,fixture', [('t1_s01_c05_ij-l_precalculated.sig', 'img_t1_s01_c05_ij.npz'),
                                          ('010067_301x300-l_precalculated.sig', 'img_010067_301x300.npz')])
def test_sig_read_and_write_back(name, fixture, tmp_path):
    import wndcharm
    data, path = raw(name, tmp_path)
    t = wndcharm.read_sig(path)
    want = load_npz(fixture)
    assert np.array_equal(t.values[0], want['values'])
    assert t.names == list(want['names'])
    assert (t.version, t.feature_vec_type, int(t.classes[0])) == (4, 2, 0)
    out = tmp_path / 'out.sig'
    wndcharm.write_sig(str(out), t.values[0], t.names, image_path=t.paths[0], sample_class=0, version=4, vec_type=2)
    assert out.read_bytes() == data


def test_sig_without_names_and_continuous(tmp_path):
    import wndcharm
    p = str(tmp_path / 'a.sig')
    wndcharm.write_sig(p, [1.0, 0.5, 1e-9], None, image_path='x.tif', sample_value=2.5, version=4, vec_type=1)
    assert open(p).read() == '2.500000\t4.1\nx.tif\n1\n0.5\n1e-09\n'
    t = wndcharm.read_sig(p, continuous=True)
    assert np.array_equal(t.values[0], [1.0, 0.5, 1e-9]) and t.sample_values[0] == 2.5 and t.classes[0] == 1


def test_fit_read_and_write_back(tmp_path):
    import wndcharm
    data, path = raw('unbalanced.fit', tmp_path)
    t = wndcharm.read_fit(path)
    want = load_npz('unbalanced_fit.npz')
    assert np.array_equal(t.values, want['values'])
    assert np.array_equal(t.classes, want['classes'])
    assert t.names == list(want['names']) and t.paths == list(want['paths'])
    assert (t.n_classes, t.version, t.feature_vec_type) == (2, 4, 2)
    assert t.class_labels == ['', 'binuc', 'norm']
    out = tmp_path / 'out.fit'
    wndcharm.write_fit(str(out), t.values, t.classes, t.names, t.class_labels, t.paths, version=4, vec_type=2)
    assert out.read_bytes() == data


def test_fit_older_writer_reads(tmp_path):
    import wndcharm
    _, path = raw('test-l.fit', tmp_path)
    t = wndcharm.read_fit(path)
    want = load_npz('test_l_fit.npz')
    assert np.array_equal(t.values, want['values'])
    assert np.array_equal(t.classes, want['classes'])
    assert t.class_labels[1:] == list(want['class_names'])


def test_weights_read_and_write_back(tmp_path):
    import wndcharm
    data, path = raw('test_fit-l.weights', tmp_path)
    w = wndcharm.read_weights(path)
    assert np.array_equal(w, load_npz('test_l_fit.npz')['weights_f1'])
    names = [ln.split(' ', 1)[1] for ln in data.decode().splitlines()]
    out = tmp_path / 'out.weights'
    wndcharm.write_weights(str(out), w, names)
    assert out.read_bytes() == data


@pytest.mark.parametrize('name,fixture,index', [('t1_s01_c05_ij.tif', 'img_t1_s01_c05_ij.npz', None),
                                                ('binuc1.tiff', 'unbalanced_fit.npz', 0),
                                                (LYMPHOMA, 'img_lymphoma_tile.npz', None)])
def test_reference_tiffs(name, fixture, index, tmp_path):
    import wndcharm
    _, path = raw(name, tmp_path)
    got = wndcharm.read_tiff(path)
    want = load_npz(fixture)['pixels']
    if index is not None:
        want = want[index]
    assert got.dtype == (np.uint8 if want.dtype == np.uint8 else np.uint16)
    assert np.array_equal(got, want)


def test_reference_tiff_8bit_strips(tmp_path):
    Image = pytest.importorskip('PIL.Image')
    import wndcharm
    _, path = raw('test-0032-0008-0008.tif', tmp_path)
    assert np.array_equal(wndcharm.read_tiff(path), np.array(Image.open(path)))


@pytest.mark.parametrize('comp', ['raw', 'tiff_lzw', 'packbits', 'tiff_adobe_deflate'])
@pytest.mark.parametrize('kind', ['u8', 'u16', 'rgb8'])
def test_pil_written_tiffs(comp, kind, tmp_path):
    Image = pytest.importorskip('PIL.Image')
    import wndcharm
    rng = np.random.default_rng(zlib.crc32((comp + kind).encode()))
    if kind == 'u8':
        a = rng.integers(0, 256, (37, 53), dtype=np.uint8)
    elif kind == 'u16':
        a = (rng.integers(0, 2000, (37, 53)) * 31).astype(np.uint16)
    else:
        # smooth image so that LZW builds long strings and wide codes
        yy, xx = np.mgrid[0:90, 0:120]
        a = np.stack([(xx * 2) % 256, (yy * 3) % 256, ((xx + yy) // 4) % 256], -1).astype(np.uint8)
    p = str(tmp_path / 'x.tif')
    info = {278: 7}  # several strips
    Image.fromarray(a).save(p, compression=comp, tiffinfo=info)
    assert np.array_equal(wndcharm.read_tiff(p), np.array(Image.open(p)))
    assert np.array_equal(wndcharm.read_tiff(p), a)


def test_predictor_two(tmp_path):
    Image = pytest.importorskip('PIL.Image')
    import wndcharm
    yy, xx = np.mgrid[0:64, 0:200]
    a = ((xx * 7 + yy * 3) % 256).astype(np.uint8)
    p = str(tmp_path / 'p.tif')
    Image.fromarray(a).save(p, compression='tiff_lzw', tiffinfo={317: 2})
    if Image.open(p).tag_v2.get(317) != 2:
        pytest.skip('this PIL does not write predictor 2')
    assert np.array_equal(wndcharm.read_tiff(p), a)


def write_rgb16_be(path, a, rows_per_strip=None):
    """a minimal big-endian uncompressed RGB 16-bit TIFF (one strip)"""
    h, w, _ = a.shape
    pix = a.astype('>u2').tobytes()
    rps = h if rows_per_strip is None else rows_per_strip
    entries = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 3, None), (259, 3, 1, 1), (262, 3, 1, 2),
               (273, 4, 1, None), (277, 3, 1, 3), (278, 4, 1, rps), (279, 4, 1, len(pix))]
    ifd_off = 8
    ifd_len = 2 + 12 * len(entries) + 4
    bps_off = ifd_off + ifd_len
    pix_off = bps_off + 6
    out = bytearray(b'MM' + struct.pack('>HI', 42, ifd_off))
    out += struct.pack('>H', len(entries))
    for tag, typ, cnt, val in entries:
        if tag == 258:
            out += struct.pack('>HHII', tag, typ, cnt, bps_off)
        elif tag == 273:
            out += struct.pack('>HHII', tag, typ, cnt, pix_off)
        elif typ == 3:
            out += struct.pack('>HHIHH', tag, typ, cnt, val, 0)
        else:
            out += struct.pack('>HHII', tag, typ, cnt, val)
    out += struct.pack('>I', 0)
    out += struct.pack('>HHH', 16, 16, 16)
    out += pix
    open(path, 'wb').write(bytes(out))


def test_rgb16_big_endian(tmp_path):
    import wndcharm
    a = np.random.default_rng(4).integers(0, 65536, (23, 31, 3)).astype(np.uint16)
    p = str(tmp_path / 'c.tif')
    write_rgb16_be(p, a)
    got = wndcharm.read_tiff(p)
    assert got.dtype == np.uint16 and np.array_equal(got, a)


def test_rows_per_strip_zero(tmp_path):
    """RowsPerStrip = 0 reads as one strip of the whole image (the field's
    default), not as an empty strip"""
    import wndcharm
    a = np.random.default_rng(5).integers(0, 65536, (9, 14, 3)).astype(np.uint16)
    p = str(tmp_path / 'z.tif')
    write_rgb16_be(p, a, rows_per_strip=0)
    got = wndcharm.read_tiff(p)
    assert np.array_equal(got, a)


def test_refused_tiffs(tmp_path):
    Image = pytest.importorskip('PIL.Image')
    import wndcharm
    p = str(tmp_path / 'f.tif')
    Image.fromarray(np.zeros((20, 20), np.float32)).save(p)
    with pytest.raises(wndcharm.WndGpuError, match='bits per sample'):
        wndcharm.read_tiff(p)
    p2 = str(tmp_path / 'm.tif')
    frames = [Image.fromarray(np.full((20, 20), k, np.uint8)) for k in range(2)]
    frames[0].save(p2, save_all=True, append_images=frames[1:])
    with pytest.raises(wndcharm.WndGpuError, match='more than one directory'):
        wndcharm.read_tiff(p2)
    with pytest.raises(wndcharm.WndGpuError, match='cannot read'):
        wndcharm.read_tiff(str(tmp_path / 'missing.tif'))
    bad = tmp_path / 'b.tif'
    bad.write_bytes(b'not a tiff at all')
    with pytest.raises(wndcharm.WndGpuError, match='not a TIFF'):
        wndcharm.read_tiff(str(bad))


def test_lzw_long_strip(tmp_path):
    """one 600 KB strip: the code width reaches 12 bits and the table is
    cleared many times"""
    Image = pytest.importorskip('PIL.Image')
    import wndcharm
    rng = np.random.default_rng(9)
    yy, xx = np.mgrid[0:600, 0:1000]
    a = np.where(rng.random((600, 1000)) < 0.5, rng.integers(0, 256, (600, 1000)), (xx // 3 + yy // 5) % 256)
    a = a.astype(np.uint8)
    p = str(tmp_path / 'l.tif')
    Image.fromarray(a).save(p, compression='tiff_lzw', tiffinfo={278: 600})
    assert np.array_equal(wndcharm.read_tiff(p), a)
