"""Summaries of rocprofv3 output for the committed profiles.

Two subcommands, each printing a Markdown table:

``trace <run_kernel_trace.csv>``
    dispatches grouped by kernel and launch shape (grid, block, LDS,
    scratch, VGPRs): calls, average / median / minimum duration.  Kernel
    names are shortened (namespaces and parameter lists dropped, the
    template arguments of mangled names decoded).  The first group in launch
    order is also split into its first ``--warmup`` dispatches and the rest,
    which for ``bench.py`` are the warm-up and the timed launches of the
    headline kernel.

``pmc <FETCH_SIZE csv> <WRITE_SIZE csv>``
    HBM bytes per dispatch from two ``rocprofv3 --pmc`` runs (one counter
    each): FETCH_SIZE and WRITE_SIZE in KiB, each scaled by what the same
    counter reads for a device copy of known size (bench.py --probe's 256
    MiB ``__amd_rocclr_copyBuffer``), so the factors come from the same run
    as the kernel figures.
"""

import argparse
import csv
import re
import statistics as st
import sys

_TARG = {'Lb1E': 'true', 'Lb0E': 'false'}


def short_name(name):
    """Readable kernel name: k_wdot_f64, k_gemm<true, true, 3>, ..."""
    nm = name.replace('(anonymous namespace)::', '')
    if nm.startswith('_Z'):
        m = re.search(r'\d(k_[A-Za-z0-9_]+?)(I(.*?)E)?E', nm)
        if m:
            base = m.group(1)
            if m.group(3):
                args = re.findall(r'L[a-z]\d+E', m.group(3) + 'E')
                vals = [_TARG.get(a, a[2:-1]) for a in args]
                return '{}<{}>'.format(base, ', '.join(vals))
            return base
    m = re.search(r'\b(k_[A-Za-z0-9_]+(<[^>]*>)?)', nm)
    if m:
        return m.group(1)
    return re.sub(r'^void ', '', nm.split('(')[0])[:60]


def trace(path, warmup=20):
    groups, order = {}, []
    with open(path) as f:
        for r in csv.DictReader(f):
            key = (short_name(r['Kernel_Name']), int(r['Grid_Size_X']),
                   int(r['Workgroup_Size_X']), int(r['LDS_Block_Size']),
                   int(r['Scratch_Size']), int(r['VGPR_Count']),
                   int(r['Accum_VGPR_Count']))
            ms = (int(r['End_Timestamp']) - int(r['Start_Timestamp'])) / 1e6
            if key not in groups:
                groups[key] = []
                order.append(key)
            groups[key].append(ms)
    out = ['| kernel | grid | block | LDS (B) | scratch (B/lane) | VGPR | '
           'AGPR | calls | avg ms | median ms | min ms | total ms |',
           '|---|---|---|---|---|---|---|---|---|---|---|---|']
    for key in sorted(order, key=lambda k: -sum(groups[k])):
        v = groups[key]
        out.append('| {} | {:,} | {} | {:,} | {} | {} | {} | {} | {:.4f} | '
                   '{:.4f} | {:.4f} | {:.2f} |'.format(
                       key[0], key[1], key[2], key[3], key[4], key[5], key[6],
                       len(v), st.mean(v), st.median(v), min(v), sum(v)))
    if order:
        v = groups[order[0]]
        rest = v[warmup:]
        out.append('')
        out.append('first group in launch order ({}, grid {:,}): {} calls; '
                   'first {}: avg {:.4f} ms; the other {}: avg {:.4f} ms, '
                   'median {:.4f} ms'.format(
                       order[0][0], order[0][1], len(v), min(warmup, len(v)),
                       st.mean(v[:warmup]), len(rest),
                       st.mean(rest) if rest else float('nan'),
                       st.median(rest) if rest else float('nan')))
    return '\n'.join(out)


def _counter(path):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((short_name(r['Kernel_Name']), int(r['Grid_Size']),
                         r['Counter_Name'], float(r['Counter_Value'])))
    return rows


def pmc(fetch_csv, write_csv, copy_bytes=256 << 20):
    fr, wr = _counter(fetch_csv), _counter(write_csv)
    if len(fr) != len(wr):
        raise SystemExit('the two runs dispatched different kernel lists')
    kib = copy_bytes / 1024.0
    # the calibration copy: the device copy whose WRITE_SIZE is nearest its
    # known size (host-to-device staging copies run under the same name)
    copies = [(abs(w[3] - kib), f[3], w[3]) for f, w in zip(fr, wr)
              if f[0].startswith('__amd_rocclr_copyBuffer')]
    _d, cf, cw = min(copies)
    ff, fw = kib / cf, kib / cw
    out = ['counter factors from the {:,}-byte copy: FETCH_SIZE read {:,.2f} '
           'KiB for {:,.0f} KiB moved -> x{:.5f}; WRITE_SIZE read {:,.2f} KiB '
           '-> x{:.5f}'.format(copy_bytes, cf, kib, ff, cw, fw), '',
           '| kernel | grid | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | fetch '
           '(B, scaled) | write (B, scaled) | total (B) |',
           '|---|---|---|---|---|---|---|']
    for f, w in zip(fr, wr):
        if not f[0].startswith('k_'):
            continue
        fb, wb = f[3] * 1024 * ff, w[3] * 1024 * fw
        out.append('| {} | {:,} | {:,.2f} | {:,.2f} | {:,.0f} | {:,.0f} | '
                   '{:,.0f} |'.format(f[0], f[1], f[3], w[3], fb, wb, fb + wb))
    return '\n'.join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    sub = ap.add_subparsers(dest='cmd', required=True)
    t = sub.add_parser('trace')
    t.add_argument('csv')
    t.add_argument('--warmup', type=int, default=20)
    p = sub.add_parser('pmc')
    p.add_argument('fetch_csv')
    p.add_argument('write_csv')
    a = ap.parse_args(argv)
    if a.cmd == 'trace':
        print(trace(a.csv, a.warmup))
    else:
        print(pmc(a.fetch_csv, a.write_csv))
    return 0


if __name__ == '__main__':
    sys.exit(main())
