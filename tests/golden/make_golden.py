#!/usr/bin/env python3
"""Write the golden fixtures used by tests/test_oracle.py.

* hiv_bimarg_pc.npy: the reference's only data fixture,
  examples/HIV_inference/example_bimarg_pc.npy (target bimarginals, L=93,
  q=4, float32), copied byte for byte (loaded with allow_pickle=False).
* philox4x32_10_kat.json: the Philox4x32-10 known-answer vectors published
  with the Random123 reference implementation (Salmon et al., SC'11):
  counter, key and expected output words.
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = Path(sys.argv[1]) if len(sys.argv) > 1 else Path('/root/reference')


def main():
    b = np.load(REF / 'examples' / 'HIV_inference' / 'example_bimarg_pc.npy', allow_pickle=False)
    np.save(HERE / 'hiv_bimarg_pc.npy', b)
    kat = [
        {"ctr": [0, 0, 0, 0], "key": [0, 0],
         "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
        {"ctr": [0xffffffff] * 4, "key": [0xffffffff] * 2,
         "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
        {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], "key": [0xa4093822, 0x299f31d0],
         "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]},
    ]
    (HERE / 'philox4x32_10_kat.json').write_text(json.dumps(kat, indent=1) + "\n")


if __name__ == '__main__':
    main()
