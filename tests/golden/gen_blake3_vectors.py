"""Write tests/golden/blake3_vectors.json from an independent BLAKE3.

The vectors come from LLVM's BLAKE3 (the llvm_blake3_hasher_* functions that
the installed libclang-cpp.so exports), called through ctypes.  Inputs follow
the official BLAKE3 test-vector convention: byte i of the input is i % 251.
The oracle's BLAKE3 (oracle/whir_oracle.c) is checked against this file.

Run:  python tests/golden/gen_blake3_vectors.py
"""
import ctypes
import glob
import json
import os

LENGTHS = [0, 1, 2, 3, 4, 5, 6, 7, 8, 63, 64, 65, 127, 128, 129, 192, 255, 256, 257, 320,
           384, 448, 511, 512, 513, 640, 768, 1023, 1024, 1025, 2048, 2049, 3072, 3073,
           4096, 4097, 5120, 5121, 6144, 6145, 7168, 7169, 8192, 8193, 16384, 31744, 102400]


def main():
    cands = sorted(glob.glob("/opt/rocm*/llvm/lib/libclang-cpp.so*"))
    lib = ctypes.CDLL(cands[0])
    init = lib.llvm_blake3_hasher_init
    update = lib.llvm_blake3_hasher_update
    final = lib.llvm_blake3_hasher_finalize
    update.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    final.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    out = []
    for n in LENGTHS:
        data = bytes(i % 251 for i in range(n))
        st = ctypes.create_string_buffer(8192)
        init(st)
        update(st, data, n)
        h = ctypes.create_string_buffer(32)
        final(st, h, 32)
        out.append({"input_len": n, "hash": h.raw.hex()})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "blake3_vectors.json")
    with open(path, "w") as f:
        json.dump({"source": "LLVM BLAKE3 (llvm_blake3_hasher_*) via ctypes; input byte i = i % 251",
                   "cases": out}, f, indent=1)
    print(path, len(out))


if __name__ == "__main__":
    main()
