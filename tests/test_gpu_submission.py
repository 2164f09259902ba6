"""Asynchronous submissions (tk_frame_submit / tk_submission_*), the receipt seam
of NativeImageSubmission (src/native/submission.rs:100-146) and the adapter's
Receipt (src/native/runtime/adapter.rs:17-47): several submissions in flight,
each keeping the image of its own render; is_complete never blocking; a frame
that exceeded a capacity read back as TK_ERR_OVERFLOW (Rejected)."""
import time

import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import _abi, scenes
from tileink_amd._abi import TileinkError

pytestmark = pytest.mark.gpu


def colored(rgb):
    """One scene structure in two paints: frames can be updated between them."""
    canvas = tk.Canvas(120, 80)
    canvas.push_rect(10, 10, 100, 60, 8.0, tk.Color(*rgb, 1.0))
    canvas.push_polygon([[20, 70], [60, 5], [110, 75]], tk.Color(rgb[2], rgb[0], rgb[1], 0.6))
    return canvas


def test_submissions_in_flight_keep_their_images(ctx):
    first, second = colored((0.9, 0.2, 0.1)), colored((0.1, 0.5, 0.8))
    frame = tk.Frame(ctx, first.encode())
    s1 = frame.submit()
    frame.update(second.encode())
    s2 = frame.submit()  # renders the second paint into the same device image
    other = tk.Frame(ctx, scenes.scene_set_member(200, 150, 300, 0).encode())
    s3 = other.submit()
    got3, got2, got1 = s3.readback(), s2.readback(), s1.readback()
    assert np.array_equal(got1, oracle.render_frame(first.arrays(), outputs=False)[0])
    assert np.array_equal(got2, oracle.render_frame(second.arrays(), outputs=False)[0])
    assert not np.array_equal(got1, got2)
    assert np.array_equal(got3, oracle.render_frame(scenes.scene_set_member(200, 150, 300, 0).arrays(),
                                                    outputs=False)[0])
    assert s1.is_complete() and s2.is_complete() and s3.is_complete()
    assert np.array_equal(s1.readback(), got1)  # a receipt can be read more than once


def test_is_complete_does_not_wait(ctx):
    canvas = scenes.mixed_scene_640x360()
    frame = tk.Frame(ctx, canvas.encode())
    frame.render()
    expected = frame.image()
    sub = frame.submit()
    deadline = time.time() + 30.0
    polls = 0
    while not sub.is_complete():
        polls += 1
        assert time.time() < deadline
        time.sleep(0.0005)
    assert np.array_equal(sub.readback(), expected)


def test_render_to_image_reads_back_on_demand(ctx):
    canvas = scenes.scene_set_member(96, 64, 50, 2)
    renderer = tk.NativeRenderer(tk.NativeBackend.Hip, 96, 64, context=ctx)
    submission = renderer.render_to_image(canvas)
    del canvas  # the submission holds its own frame
    image = submission.readback()
    assert image.shape == (64, 96)


def test_overflow_is_rejected_at_readback(ctx):
    canvas = scenes.mixed_scene_640x360()
    arrays = canvas.arrays()
    arrays.scalars["ptcl_capacity"] = 4  # the frame's PTCL buffer holds four records
    frame = tk.Frame(ctx, arrays.to_scene())
    sub = frame.submit()
    with pytest.raises(TileinkError) as err:
        sub.readback()
    assert err.value.status == _abi.TK_ERR_OVERFLOW
    assert frame.stats()["overflow"] & 4


def test_batch_frame_is_refused(ctx):
    batch = tk.Batch(ctx, [scenes.scene_set_member(64, 64, 20, 3).encode()])
    with pytest.raises(TileinkError):
        batch.frame.submit()
