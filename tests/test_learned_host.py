"""Host-side logic of gaukuk_amd.learned that needs no GPU."""
import numpy as np


def test_continued_finetuning_record():
    """A fine-tuning run on a network fine-tuned before adds up the iterations, appends its history
    after the earlier one (offset by the earlier total) and drops the earlier run's settings; on a
    network never fine-tuned it starts the record."""
    from gaukuk_amd.learned import continued_record
    first = continued_record({"seeds": np.arange(3)}, [[0, 0.9], [5, 0.8]], 10)
    assert first["ft_iters"] == 10 and np.array_equal(first["seeds"], np.arange(3))
    assert np.array_equal(first["ft_history"], [[0, 0.9], [5, 0.8]])
    prev = dict(first, ft_sigma=0.02, ft_lr=0.01)
    second = continued_record(prev, [[0, 0.79], [5, 0.7]], 20)
    assert second["ft_iters"] == 30
    assert np.array_equal(second["ft_history"], [[0, 0.9], [5, 0.8], [10, 0.79], [15, 0.7]])
    assert "ft_sigma" not in second and "ft_lr" not in second
    assert np.array_equal(second["seeds"], np.arange(3))
