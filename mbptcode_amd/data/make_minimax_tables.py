"""Regenerate ``minimax_tau.json`` from the GreenX minimax imaginary-time tables.

The numbers are the GreenX "Laplace" coefficients (GX-TimeFrequency,
minimax_tau.F90; Apache-2.0, see ``NOTICE`` beside this file) as the reference
ships them in ``src/Base/utils/grids/minimax_tau_data.json``.  They are data,
not code: a minimax grid is a table lookup, so the product needs the same
numbers.  This script only re-serialises them into this package's layout

    {"<ntau>": {"ranges": [R_1, ..., R_n], "rows": [[tau_1..tau_n, w_1..w_n], ...]}}

where row j serves e_max/e_min < ranges[j] and the last row everything wider.
It runs only where the reference tree is mounted; the JSON it writes is
committed, so nothing at run time reads the reference.
"""
import json
import os
import sys

SRC = '/root/reference/src/Base/utils/grids/minimax_tau_data.json'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'minimax_tau.json')


def main(src=SRC, out=OUT):
    with open(src) as f:
        table = json.load(f)
    packed = {}
    for size in sorted(table, key=int):
        entry = table[size]
        packed[size] = {'ranges': [float(x) for x in entry['energy_range']],
                        'rows': [[float(x) for x in row]
                                 for row in entry['aw_erange_matrix']]}
    with open(out, 'w') as f:
        json.dump({'origin': 'GreenX GX-TimeFrequency minimax_tau (Apache-2.0)',
                   'tables': packed}, f, separators=(',', ':'))
    return out


if __name__ == '__main__':
    print(main(*sys.argv[1:]))
