"""tcv: MI355X-native W-projection gridding / degridding behind tConvolve's
Benchmark surface.  All compute goes through libtcv.so (HIP, gfx950)."""
from ._lib import LIB_PATH, exported_symbols, lib
from .benchmark import Benchmark, CurrentBenchmark
from .device import (Batch, Context, DeviceArray, HalfTable, HalfVis, Plan, Server, TcvError,
                     Visibilities, attic_initC_device, autotune, degrid_plain, degrid_plain_f64,
                     grid_plain, grid_plain_f64)
from .gridder import GridJob, init_offsets
from . import setup

__all__ = ["autotune", "Server", "Batch", "attic_initC_device", "LIB_PATH", "exported_symbols", "lib", "Benchmark", "CurrentBenchmark", "Context",
           "DeviceArray", "HalfTable", "HalfVis", "Plan", "TcvError", "Visibilities", "degrid_plain", "grid_plain",
           "degrid_plain_f64", "grid_plain_f64",
           "GridJob", "init_offsets", "setup"]
