"""Golden vectors for the order-1 (bilinear) warps of the off-axis / laser-guide-star footprints.

The reference interpolates with scikit-image (tools/interpolateGeometricalTransformation.py:
interpolate_image / interpolate_cube: a composition of SimilarityTransforms passed as
warp(image, transform.inverse, output_shape, order=1), mode 'constant', cval 0). This script
RUNS scikit-image on seeded synthetic arrays for the two compositions those functions reduce
to here -- a pure translation (ratio 1) and a scale by ratio about the image centre followed by
the translation (extra, extra) -- and stores inputs, parameters and outputs. It reads no
reference source.

Run:  /opt/conda/bin/python3.9 tests/golden/make_bilinear_golden.py
Writes tests/golden/bilinear_skimage.npz
"""
import os

import numpy as np
import skimage
import skimage.transform as sk

HERE = os.path.dirname(os.path.abspath(__file__))


def scale_about_centre(shape, ratio):
    sy, sx = np.array(shape[:2]) / 2.0
    return (sk.SimilarityTransform(translation=[-sx, -sy]) + (sk.SimilarityTransform(scale=ratio)
                                                               + sk.SimilarityTransform(translation=[sx, sy])))


def main():
    rs = np.random.RandomState(4321)
    cases = {}
    k = 0
    for shape in [(12, 12), (16, 16), (9, 9)]:
        img = np.cumsum(np.cumsum(rs.randn(*shape), 0), 1) * 0.1
        for sx, sy in [(0.0, 0.0), (-0.3, 0.0), (0.0, -0.7), (-0.25, -0.6), (0.4, 0.55), (-0.999, 0.001)]:
            tf = sk.SimilarityTransform(translation=[sx, sy])
            cases['img_%d' % k] = img
            cases['kind_%d' % k] = np.array(0)
            cases['par_%d' % k] = np.array([sx, sy])
            cases['out_%d' % k] = sk.warp(img, tf.inverse, output_shape=shape, order=1)
            k += 1
        for m in [0.95, 0.8, 0.5, 0.9999]:
            ratio = 1.0 / m
            extra = (ratio % 1) / 2 + (np.floor(ratio) - 1) * 0.5
            tf = scale_about_centre(shape, ratio) + sk.SimilarityTransform(translation=[extra, extra])
            cases['img_%d' % k] = img
            cases['kind_%d' % k] = np.array(1)
            cases['par_%d' % k] = np.array([ratio, extra])
            cases['out_%d' % k] = sk.warp(img, tf.inverse, output_shape=shape, order=1)
            k += 1
    cases['n_cases'] = np.array(k)
    cases['skimage_version'] = np.array(skimage.__version__)
    np.savez(os.path.join(HERE, 'bilinear_skimage.npz'), **cases)
    print('wrote', k, 'cases with scikit-image', skimage.__version__)


if __name__ == '__main__':
    main()
