"""MI355X FE part reducer: Python side of libfedem_reducer_core.so.

``gpu``      ctypes bindings of the plug-in seam (include/fedem_reducer_gpu.h)
``parts``    seeded synthetic FE parts (plates, HEX8 blocks, beam frames)
``reducer``  FedemReducer, the fedempy-compatible driver of the drop-in ABI
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(ROOT_DIR, "lib", "libfedem_reducer_core.so")
