"""The Python boundary rejects inputs the C ABI would read or write through a
bad pointer: host tensors, wrong dtypes, undersized images, a light list
without its count, and a multi-job tag that is still in flight.  No GPU call
is made: every case fails in the argument checks."""
import numpy as np
import pytest
import torch

from mambatracer_amd import api
from mambatracer_amd.workloads import RenderJob


def _job(w=8, h=4):
    return RenderJob(width=w, height=h, spp=1, max_bounces=2, rr_start=-1, seed=1)


def test_materials_must_be_device_float32_records():
    job = _job()
    with pytest.raises(ValueError, match="materials"):
        api._check_job_inputs(job, torch.zeros((3, 36), dtype=torch.float32))      # host tensor
    with pytest.raises(ValueError, match="materials"):
        api._check_job_inputs(job, np.zeros((3, 36), np.float32))                  # not a tensor


def test_light_list_goes_with_its_count():
    job = _job()
    with pytest.raises(ValueError):
        api._check_job_inputs(job, None, light_faces=torch.zeros(3, dtype=torch.int32))


def test_image_shape_is_checked():
    job = _job()
    with pytest.raises(ValueError, match="image"):
        api._need(torch.zeros((job.width * job.height - 1, 4)), "image", job.width * job.height, (4,))


def test_batch_rejects_a_tag_in_flight():
    b = api.Batch.__new__(api.Batch)  # no device engine: the tag check comes first
    b._keep = {7: ()}
    b.handle = None
    with pytest.raises(ValueError, match="in flight"):
        b.submit(7, None, None, None, _job(), None)
