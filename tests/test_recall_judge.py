"""The recall judge (compactcnncascade_amd/recall.py) against the closed form of the
reference's Rect::overlap (type.h:154-173) and hand-checked pairings. CPU only."""
import numpy as np

from compactcnncascade_amd.recall import match, recall, rect_overlap


def test_overlap_closed_form():
    assert rect_overlap((0, 0, 10, 10), (0, 0, 10, 10)) == 1.0
    # half shifted: intersection 5x10 over the 15x10 bounding box
    assert abs(rect_overlap((0, 0, 10, 10), (5, 0, 10, 10)) - 50 / 150) < 1e-12
    # touching edges count as touching (the '>' test) with zero intersection
    assert rect_overlap((0, 0, 10, 10), (10, 0, 10, 10)) == 0.0
    assert rect_overlap((0, 0, 10, 10), (11, 0, 10, 10)) == 0.0
    # a one-pixel shift of a 100-pixel box stays above 0.98 but below 0.99
    o = rect_overlap((0, 0, 100, 100), (1, 0, 100, 100))
    assert abs(o - 99 * 100 / (101 * 100)) < 1e-12 and 0.98 < o < 0.99
    # nested: inner area over the outer box
    assert abs(rect_overlap((0, 0, 100, 100), (0, 0, 100, 99)) - 0.99) < 1e-12


def test_greedy_one_to_one_pairing():
    ref = np.array([[0, 0, 100, 100], [500, 500, 60, 60]])
    # two test boxes near the first reference box: only one pairs with it
    test = np.array([[0, 0, 100, 100], [0, 0, 100, 99], [500, 500, 60, 60]])
    assert match(ref, test) == 2
    assert match(ref, test[:2]) == 1
    assert match(ref, np.zeros((0, 4))) == 0
    r, paired, nref, ntest = recall([ref, np.zeros((0, 4))], [test, np.array([[1, 1, 5, 5]])])
    assert (r, paired, nref, ntest) == (1.0, 2, 2, 4)
    assert recall([np.zeros((0, 4))], [np.zeros((0, 4))])[0] == 1.0
