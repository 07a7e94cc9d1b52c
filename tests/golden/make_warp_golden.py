"""Golden vectors for the atmosphere's sub-pixel shift (tools/tools.py:202-209).

The reference shifts its phase screens with
    skimage.transform.warp(image, SimilarityTransform(translation=shift).inverse, order=3)
(default mode='constant', cval=0, clip=True). scikit-image is a third-party
dependency (requirements.txt:26-27 pins 0.19.x; this image has 0.18.3 under
/opt/conda). This script RUNS that library on seeded synthetic arrays and
stores inputs and outputs; it reads no reference source.

Run:  /opt/conda/bin/python3.9 tests/golden/make_warp_golden.py
Writes tests/golden/warp_skimage.npz
"""
import os

import numpy as np
import skimage
import skimage.transform as sk

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rs = np.random.RandomState(1234)
    cases = {}
    shapes = [(12, 12), (17, 23)]
    shifts = [(0.0, 0.0), (0.4, 0.0), (0.0, 0.4), (0.8, 4.9e-14), (-0.3, 0.7), (1.0, 0.0), (0.0, -1.0),
              (1.0, 1.0), (0.999, -0.25), (-0.6, -0.55)]
    k = 0
    for shp in shapes:
        for offset in (0.0, 3.0):            # offset 3: all-positive image (cval outside [min, max])
            img = np.cumsum(np.cumsum(rs.randn(*shp), 0), 1) * 0.1 + offset
            for (tx, ty) in shifts:
                tf = sk.SimilarityTransform(translation=[tx, ty])
                out = sk.warp(img, tf.inverse, order=3)
                cases['img_%d' % k] = img
                cases['shift_%d' % k] = np.array([tx, ty])
                cases['out_%d' % k] = out
                k += 1
    cases['n_cases'] = np.array(k)
    cases['skimage_version'] = np.array(skimage.__version__)
    np.savez_compressed(os.path.join(HERE, 'warp_skimage.npz'), **cases)
    print('wrote %d cases with scikit-image %s' % (k, skimage.__version__))


if __name__ == '__main__':
    main()
