"""Pins of the extension-driver oracle (ExtensionAligner.extend)."""
from oracle import pyoracle as orc


def region(seeds, read, ref, **kw):
    return orc.extend_chain(seeds, read, ref, sc=orc.scoring(**kw))


def test_seed_only_score():                                 # AlignmentTests.swift:1140-1170
    r = region([(100, 0, 10, 10)], [0] * 10, [0] * 200)
    assert r[0]["score"] == 10


def test_left_only_score():                                 # :1172-1203
    r = region([(105, 5, 5, 5)], [0] * 10, [0] * 200, pen_clip5=100)
    assert r[0]["score"] == 10


def test_right_only_score():                                # :1205-1236
    r = region([(100, 0, 5, 5)], [0] * 10, [0] * 200, pen_clip3=100)
    assert r[0]["score"] == 10


def test_both_sides_score():                                # :1238-1268
    r = region([(103, 3, 4, 4)], [0] * 10, [0] * 200, pen_clip5=100, pen_clip3=100)
    assert r[0]["score"] == 10


def test_covered_seed_sub():                                # :1309-1345
    r = region([(100, 0, 30, 30), (110, 10, 20, 20)], [0] * 50, [0] * 300, pen_clip5=100,
               pen_clip3=100)
    assert len(r) == 1 and r[0]["sub"] == 20


def test_covered_seed_below_threshold():                    # :1347-1380
    r = region([(100, 0, 30, 30), (105, 5, 10, 10)], [0] * 50, [0] * 300, pen_clip5=100,
               pen_clip3=100)
    assert len(r) == 1 and r[0]["sub"] == 0


def test_clip_vs_extend():                                  # :783-826
    read = [3] * 10 + [0] * 20 + [3] * 10
    r = region([(100, 10, 20, 20)], read, [0] * 200, pen_clip5=5, pen_clip3=5)
    assert r and r[0]["qb"] >= 0 and r[0]["qe"] <= 40
