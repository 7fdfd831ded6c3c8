"""Write the golden table fixtures from the reference's generated sources.

The reference ships the generated OpenCL output of its own code generator for
two mechanisms (``examples/opencl/h2o2``: H2/O2, CONP, F order, full Jacobian;
``examples/opencl/ch4``: GRI-Mech 3.0, CONP, F order, sparse CCS Jacobian).
Their headers hold every mechanism table as a ``__constant`` array literal and
the main file holds the species names and reaction strings.  This script reads
those literals as data (numbers and strings only) and stores them as JSON so
the table builders can be checked against them without the reference present.

Run from the repository root:  python tests/golden/extract_reference_tables.py
"""

import json
import os
import re
import sys

REF = os.environ.get('PYJAC_REFERENCE', '/root/reference')
HERE = os.path.dirname(os.path.abspath(__file__))

ARRAY = re.compile(
    r'__constant\s+(int|double)\s+const\s+(\w+)\[([0-9 *]+)\]\s*=\s*\{([^}]*)\}')
STRINGS = re.compile(r'Kernel::_(species_names|rxn_strings)\[\]\s*=\s*\{(.*?)\};',
                     re.S)
SCALARS = re.compile(r'#define\s+(NS|NR|NN)\s+\((\d+)\)')


def extract(folder):
    out = {'arrays': {}, 'strings': {}, 'sizes': {}}
    for name in sorted(os.listdir(folder)):
        path = os.path.join(folder, name)
        if not (name.endswith('.oclh') or name.endswith('.ocl')):
            continue
        with open(path) as fh:
            text = fh.read()
        for ctype, var, size, body in ARRAY.findall(text):
            n = eval(size, {}, {})  # sizes are products of integer literals
            vals = [v.strip() for v in body.split(',') if v.strip()]
            if ctype == 'int':
                vals = [int(v) for v in vals]
            else:
                vals = [float(v) for v in vals]
            assert len(vals) == n, (var, len(vals), n)
            out['arrays'][var] = {'dtype': ctype, 'values': vals}
        for key, body in STRINGS.findall(text):
            out['strings'][key] = re.findall(r'"((?:[^"\\]|\\.)*)"', body)
        for key, val in SCALARS.findall(text):
            out['sizes'][key] = int(val)
    return out


def main():
    for src, dst in (('h2o2', 'h2o2_tables.json'), ('ch4', 'gri30_tables.json')):
        folder = os.path.join(REF, 'examples', 'opencl', src)
        if not os.path.isdir(folder):
            sys.exit('reference example folder missing: ' + folder)
        data = extract(folder)
        with open(os.path.join(HERE, dst), 'w') as fh:
            json.dump(data, fh, separators=(',', ':'))
        print(dst, len(data['arrays']), 'arrays', data['sizes'])


if __name__ == '__main__':
    main()
