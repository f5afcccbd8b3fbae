"""The learned warm start's training and held-out tones come from disjoint
seeds: batch b of 256 tones is drawn with TRAIN_SEED + b for training and
EVAL_SEED + b for evaluation (the benchmark's 4096 held-out tones are
batches 0..15; the trained-network GPU test uses EVAL_SEED + 1000)."""
import numpy as np


def test_training_and_held_out_seed_ranges_are_disjoint():
    from cochlearmodel_amd import warmstart as W
    train = set(range(W.TRAIN_SEED, W.TRAIN_SEED + W.SEED_RANGE))
    held_out = set(range(W.EVAL_SEED, W.EVAL_SEED + 1001))
    assert not (train & held_out)
    meta = W.load_model()[2]
    if "seed" in meta:   # the committed network was trained on the training range
        n_batches = int(meta["n_signals"]) // 256
        assert W.TRAIN_SEED <= int(meta["seed"]) and int(meta["seed"]) + n_batches <= W.TRAIN_SEED + W.SEED_RANGE


def test_held_out_tones_differ_from_training_tones():
    from cochlearmodel_amd import batch, warmstart as W
    ft, _ = batch.random_tones(256, W.TRAIN_SEED)
    fe, _ = batch.random_tones(256, W.EVAL_SEED)
    assert not np.intersect1d(ft, fe).size
