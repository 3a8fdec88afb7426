"""ggks-multiselect_amd: MI355X-native selection (bucket multiselect, MSD radix
multiselect, top-K) behind the C ABI of libggks.so.

The directory name holds a dash, so it is imported by path:
``load_package()`` in __graft_entry__.py registers it as
``ggks_multiselect_amd``.
"""
from .api import (Batch, Context, DeviceArray, DeviceView, GGKSError,  # noqa: F401
                  RadixSelectWrapper, bucketMultiselectWrapper, generate, multiselect,
                  multiselect_async, randomizedTopkSelectWrapper, scratch_bytes,
                  segment_table, Server, sync, topk, tuning_get, tuning_reset, tuning_set,
                  tuning_single_get, tuning_single_set)
from . import workloads  # noqa: F401
from . import serving  # noqa: F401
from . import autotune  # noqa: F401
from . import compare  # noqa: F401

__version__ = "0.1"
