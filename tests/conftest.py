"""Shared test setup: the `gpu` marker, the repository on sys.path, scene fixtures."""

from __future__ import annotations

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
  sys.path.insert(0, str(ROOT))

# Exposure 1: linear HDR is display-relative (1e-3 absolute tolerances are meaningful).
EV100_UNIT = float(__import__("math").log2(1.0 / 1.2))


def pytest_configure(config):
  config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
  config.addinivalue_line("markers", "slow: long-running (high sample counts)")


def gpu_available() -> bool:
  import torch

  return torch.cuda.is_available()


@pytest.fixture(scope="session")
def cuda():
  import torch

  if not torch.cuda.is_available():
    pytest.skip("no GPU")
  return torch.device("cuda:0")
