"""MI355X-native TiltRec reconstruction path (AuTom2).

The product is libtiltrec_hip.so (C ABI in include/tiltrec.h) plus the
TiltRec-compatible CLI (autom2_amd/bin/tiltrec); `autom2_amd.tiltrec` is the
Python binding of the library.
"""
import os

ROOT = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(ROOT, "lib")
BIN_DIR = os.path.join(ROOT, "bin")
