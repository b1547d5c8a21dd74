"""Regenerate ``minimax_tau.json`` and ``minimax_omega.json`` from the GreenX
minimax tables.

The numbers are the GreenX coefficients (GX-TimeFrequency, minimax_tau.F90 for
the imaginary-time "Laplace" grids and minimax_omega.F90 for the
imaginary-frequency grids; Apache-2.0, see ``NOTICE`` beside the tables in
mbptcode_amd/data/) as the
reference ships them in ``src/Base/utils/grids/minimax_tau_data.json`` and
``minimax_omega_data.json``.  They are data, not code: a minimax grid is a
table lookup, so the product needs the same numbers.  This script only
re-serialises them into this package's layout

    {"<n>": {"ranges": [R_1, ..., R_n], "rows": [[x_1..x_n, w_1..w_n], ...]}}

where row j serves e_max/e_min < ranges[j] and the last row everything wider.
It runs only where the reference tree is mounted; the JSON it writes is
committed, so nothing at run time reads the reference.
"""
import json
import os
import sys

SRC_DIR = os.path.join(os.environ.get('MBPT_REFERENCE', '/root/reference'),
                       'src', 'Base', 'utils', 'grids')
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), '..', '..', 'mbptcode_amd',
                    'data')
TABLES = (('minimax_tau_data.json', 'minimax_tau.json',
           'GreenX GX-TimeFrequency minimax_tau (Apache-2.0)'),
          ('minimax_omega_data.json', 'minimax_omega.json',
           'GreenX GX-TimeFrequency minimax_omega (Apache-2.0)'))


def pack(src, out, origin):
    with open(src) as f:
        table = json.load(f)
    packed = {}
    for size in sorted(table, key=int):
        entry = table[size]
        packed[size] = {'ranges': [float(x) for x in entry['energy_range']],
                        'rows': [[float(x) for x in row]
                                 for row in entry['aw_erange_matrix']]}
    with open(out, 'w') as f:
        json.dump({'origin': origin, 'tables': packed}, f, separators=(',', ':'))
    return out


def main(src_dir=SRC_DIR, out_dir=DATA_DIR):
    return [pack(os.path.join(src_dir, s), os.path.join(out_dir, o), origin)
            for s, o, origin in TABLES]


if __name__ == '__main__':
    print(main(*sys.argv[1:]))
