"""Writes the golden fixtures under tests/golden/ from data files that the
reference's own tests hold (TIFF inputs, .sig/.fit/.weights outputs and the
expected WND classification table of tests/wndchrm_tests/test01-classify.pl).

Run once in a container that has /root/reference; the fixtures are committed.
Only data is copied: pixel arrays, feature values and names, class labels.
"""
import gzip
import os
import re
import numpy as np
from PIL import Image

REF = '/root/reference/tests'
OUT = os.path.dirname(os.path.abspath(__file__))


def read_sig(path):
    lines = open(path).read().splitlines()
    vals, names = [], []
    for ln in lines[2:]:
        v, n = ln.split('\t', 1)
        vals.append(float(v))
        names.append(n)
    return np.array(vals), names


def read_fit(path):
    """TrainingSet::ReadFromFile layout (src/wndchrm_src/TrainingSet.cpp:388-502)."""
    lines = open(path).read().splitlines()
    n_classes = int(lines[0].split()[0])
    nf = int(lines[1])
    ns = int(lines[2])
    names = lines[3:3 + nf]
    pos = 3 + nf
    assert lines[pos] == ''
    class_names = lines[pos + 1:pos + 1 + n_classes]
    pos += 1 + n_classes
    data, cls, paths = [], [], []
    for s in range(ns):
        toks = lines[pos].split()
        data.append([float(t) for t in toks[:nf]])
        cls.append(int(toks[nf]))
        paths.append(lines[pos + 1])
        pos += 2
    return np.array(data), np.array(cls, np.int32), names, class_names, paths


def load_tiff(path):
    return np.array(Image.open(path))


def main():
    w = os.path.join(REF, 'wndchrm_tests')
    p = os.path.join(REF, 'pywndcharm_tests')
    sig_cases = [
        ('img_010067_301x300', os.path.join(w, '010067_301x300.tif'), os.path.join(w, '010067_301x300-l_precalculated.sig')),
        ('img_t1_s01_c05_ij', os.path.join(w, 't1_s01_c05_ij.tif'), os.path.join(w, 't1_s01_c05_ij-l_precalculated.sig')),
        ('img_lymphoma_tile', os.path.join(p, 'lymphoma_eosin_channel_MCL_test_img_sj-05-3362-R2_001_E-t6x5_5_4-l.tiff'),
         os.path.join(p, 'lymphoma_eosin_channel_MCL_test_img_sj-05-3362-R2_001_E-t6x5_5_4-l.sig')),
    ]
    for name, tif, sig in sig_cases:
        vals, names = read_sig(sig)
        np.savez_compressed(os.path.join(OUT, name + '.npz'), pixels=load_tiff(tif), values=vals,
                            names=np.array(names))
    # unbalanced.fit: 8 images x 2919 features, exact-diff golden of test04-sigs.pl
    data, cls, names, cn, paths = read_fit(os.path.join(w, 'unbalanced.fit'))
    imgs = []
    for pth in paths:
        base = os.path.basename(pth)
        sub = 'binuc' if base.startswith('binuc') else 'norm'
        imgs.append(load_tiff(os.path.join(w, 'images', sub, base)))
    np.savez_compressed(os.path.join(OUT, 'unbalanced_fit.npz'), pixels=np.stack(imgs), values=data,
                        classes=cls, names=np.array(names), paths=np.array(paths))
    # test-l.fit + Fisher weights (-f1.0) + expected classify table
    data, cls, names, cn, paths = read_fit(os.path.join(w, 'test-l.fit'))
    wts = [float(l.split(' ', 1)[0]) for l in open(os.path.join(w, 'test_fit-l.weights')).read().splitlines() if l.strip()]
    table = re.search(r"<<END;\n(.*?)\nEND", open(os.path.join(w, 'test01-classify.pl')).read(), re.S).group(1)
    rows = [r.split('\t') for r in table.splitlines()]
    np.savez_compressed(os.path.join(OUT, 'test_l_fit.npz'), values=data, classes=cls, names=np.array(names),
                        class_names=np.array(cn), weights_f1=np.array(wts),
                        kat_norm=np.array([float(r[1]) for r in rows]),
                        kat_p=np.array([[float(r[2]), float(r[3])] for r in rows]),
                        kat_pred=np.array([cn.index(r[5]) + 1 for r in rows], np.int32),
                        kat_interp=np.array([float(r[6]) for r in rows]))
    write_raw(w, p)


def write_raw(w=os.path.join(REF, 'wndchrm_tests'), p=os.path.join(REF, 'pywndcharm_tests')):
    # data files held verbatim (gzip-compressed) for the file-format tests:
    # .sig / .fit / weights text as the reference wrote it, and TIFF inputs
    # (uncompressed big-endian 16-bit, LZW 16-bit, 8-bit strips; the eight
    # images of unbalanced.fit for the end-to-end .fit test)
    raw = os.path.join(OUT, 'raw')
    os.makedirs(raw, exist_ok=True)
    for src in [os.path.join(w, 't1_s01_c05_ij-l_precalculated.sig'),
                os.path.join(w, '010067_301x300-l_precalculated.sig'),
                os.path.join(w, 'unbalanced.fit'), os.path.join(w, 'test-l.fit'),
                os.path.join(w, 'test_fit-l.weights'), os.path.join(w, 'test_fit-l-normalized.fit'),
                os.path.join(w, 't1_s01_c05_ij.tif'),
                os.path.join(w, 'images', 'binuc', 'binuc1.tiff'),
                os.path.join(w, 'images', 'binuc', 'binuc2.tiff'),
                os.path.join(w, 'images', 'binuc', 'binuc3.tiff'),
                os.path.join(w, 'images', 'binuc', 'binuc4.tiff'),
                os.path.join(w, 'images', 'binuc', 'binuc5.tiff'),
                os.path.join(w, 'images', 'norm', 'norm1.tiff'),
                os.path.join(w, 'images', 'norm', 'norm2.tiff'),
                os.path.join(w, 'images', 'norm', 'norm3.tiff'),
                os.path.join(p, 'lymphoma_eosin_channel_MCL_test_img_sj-05-3362-R2_001_E-t6x5_5_4-l.tiff'),
                os.path.join(p, 'test-0032-0008-0008.tif')]:
        data = open(src, 'rb').read()
        with gzip.GzipFile(os.path.join(raw, os.path.basename(src) + '.gz'), 'wb', mtime=0) as f:
            f.write(data)


if __name__ == '__main__':
    main()
