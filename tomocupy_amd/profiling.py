"""Summaries of rocprofv3 output for this library's kernels.

Usage (paths are rocprofv3 ``-f csv`` outputs)::

  python -m tomocupy_amd.profiling stats  KERNEL_STATS_CSV  [--title T] [--cmd C]
  python -m tomocupy_amd.profiling traffic FETCH_COUNTER_CSV WRITE_COUNTER_CSV
         --copy-bytes B [--key-suffix S] [--json OUT.json] [--title T] [--cmd C]
  python -m tomocupy_amd.profiling window KERNEL_TRACE_CSV --kernel K --skip N --count M
  python -m tomocupy_amd.profiling counters COUNTER_CSV [COUNTER_CSV ...] [--kernels K1,K2]
         [--title T] [--cmd C]

``stats`` prints a markdown table of the kernel-trace statistics.
``counters`` averages every ``--pmc`` counter per dispatch and kernel over
one or more counter-collection files (one per pass) and prints the raw
averages plus per-wave and utilisation figures derived from them.  ``window``
prints the mean duration of dispatches N .. N+M-1 (in start order) of the
kernels whose base name is K, e.g. the launches of one pass of a run.  ``traffic``
averages ``FETCH_SIZE`` and ``WRITE_SIZE`` (KiB per dispatch) per kernel from
two ``--pmc`` runs and scales each by the factor that makes the calibration
copy kernel (the largest device-to-device copy in the trace, ``--copy-bytes``
read and written) read and write exactly its size.
"""
import argparse
import csv
import json
import re
import sys
from collections import defaultdict


_MANGLED_ARGS = (('DF16_', '_Float16'), ('DF16b', '__bf16'), ('f', 'float'), ('d', 'double'))


def _demangle_kernel(name):
    """``k_name<args>`` from an Itanium-mangled kernel symbol that the trace
    left mangled (its demangler does not know _Float16 / __bf16)."""
    pos, ident = 3 if name.startswith('_ZN') else 2, None
    while pos < len(name):  # <length><identifier> components
        m = re.match(r'\d+', name[pos:])
        if not m:
            return None
        ln = int(m.group(0))
        pos += m.end()
        comp = name[pos:pos + ln]
        pos += ln
        if comp.startswith('k_'):
            ident = comp
            break
    if ident is None:
        return None
    rest = name[pos:]
    if not rest.startswith('I'):
        return ident
    rest, args = rest[1:], []
    while rest and not rest.startswith('E'):
        lit = re.match(r'Li(-?\d+)E', rest)
        if lit:
            args.append(lit.group(1))
            rest = rest[lit.end():]
            continue
        for code, text in _MANGLED_ARGS:
            if rest.startswith(code):
                args.append(text)
                rest = rest[len(code):]
                break
        else:
            return ident
    return '%s<%s>' % (ident, ', '.join(args))


def short_name(name):
    if name.startswith('_Z'):
        d = _demangle_kernel(name)
        if d:
            return d
    m = re.search(r'(k_[A-Za-z0-9_]+)', name)
    if m:
        base = m.group(1)
        t = re.search(r'k_[A-Za-z0-9_]+<([^>]*)>', name)
        return base + ('<%s>' % t.group(1) if t else '')
    if 'copyBuffer' in name or 'direct_copy' in name or 'elementwise_kernel' in name:
        return 'copy:' + name[:40]
    return name[:60]


def cmd_stats(a):
    rows = list(csv.DictReader(open(a.csv)))
    out = []
    if a.title:
        out.append('# %s\n' % a.title)
    if a.cmd:
        out.append('Command:\n\n```\n%s\n```\n' % a.cmd)
    out.append('| kernel | calls | total ms | avg us | min us | max us | % |')
    out.append('|---|---|---|---|---|---|---|')
    for r in rows:
        name = short_name(r['Name'])
        if not name.startswith('k_'):
            continue
        out.append('| %s | %s | %.3f | %.1f | %.1f | %.1f | %.2f |' % (
            name, r['Calls'], float(r['TotalDurationNs']) / 1e6, float(r['AverageNs']) / 1e3,
            float(r['MinNs']) / 1e3, float(r['MaxNs']) / 1e3, float(r['Percentage'])))
    print('\n'.join(out))


def cmd_window(a):
    rows = [r for r in csv.DictReader(open(a.csv))
            if short_name(r['Kernel_Name']).split('<')[0] == a.kernel]
    rows.sort(key=lambda r: int(r['Start_Timestamp']))
    sel = rows[a.skip:a.skip + a.count]
    if not sel:
        sys.exit('no dispatches of %s in [%d, %d)' % (a.kernel, a.skip, a.skip + a.count))
    d = [(int(r['End_Timestamp']) - int(r['Start_Timestamp'])) / 1e3 for r in sel]
    print('%s: dispatches %d..%d of %d, mean %.1f us, min %.1f us, max %.1f us'
          % (a.kernel, a.skip, a.skip + len(sel) - 1, len(rows), sum(d) / len(d), min(d), max(d)))


def _per_kernel(path, counter):
    acc = defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r.get('Counter_Name', counter) != counter:
            continue
        acc[r['Kernel_Name']].append(float(r['Counter_Value']))
    return acc


def cmd_traffic(a):
    fetch = _per_kernel(a.fetch, 'FETCH_SIZE')
    write = _per_kernel(a.write, 'WRITE_SIZE')
    # calibration: the dispatch with the largest fetch among copy kernels
    copies = [(max(v), k) for k, v in fetch.items() if short_name(k).startswith('copy:')]
    if not copies:
        sys.exit('no copy kernel found for calibration')
    cf, ck = max(copies)
    cw = max(write.get(ck, [0.0]))
    f_scale = a.copy_bytes / (cf * 1024.0)
    w_scale = a.copy_bytes / (cw * 1024.0) if cw else 1.0
    res = {}
    base_acc = defaultdict(lambda: [0.0, 0])  # base name -> [bytes summed over dispatches, count]
    lines = []
    if a.title:
        lines.append('# %s\n' % a.title)
    if a.cmd:
        lines.append('Commands:\n\n```\n%s\n```\n' % a.cmd)
    lines.append('Calibration: %s copied %.1f MB; raw FETCH_SIZE %.1f MB, WRITE_SIZE %.1f MB '
                 '-> scale %.3f (fetch), %.3f (write).\n'
                 % (short_name(ck), a.copy_bytes / 1e6, cf * 1024 / 1e6, cw * 1024 / 1e6,
                    f_scale, w_scale))
    lines.append('| kernel | dispatches | fetch MB/dispatch | write MB/dispatch | total MB |')
    lines.append('|---|---|---|---|---|')
    for k in sorted(fetch):
        sn = short_name(k)
        if not sn.startswith('k_'):
            continue
        fv = fetch[k]
        wv = write.get(k, [0.0])
        fb = sum(fv) / len(fv) * 1024 * f_scale
        wb = sum(wv) / len(wv) * 1024 * w_scale
        base = sn.split('<')[0]
        base_acc[base][0] += (fb + wb) * len(fv)
        base_acc[base][1] += len(fv)
        lines.append('| %s | %d | %.2f | %.2f | %.2f |' % (sn, len(fv), fb / 1e6, wb / 1e6,
                                                          (fb + wb) / 1e6))
    # per base kernel name (all template instances and launch shapes): mean
    # HBM bytes per dispatch, the figure the bench's roofline pairs with the
    # mean algorithmic bytes per launch
    for base, (tot, cnt) in base_acc.items():
        res[base + (a.key_suffix or '')] = round(tot / cnt)
    print('\n'.join(lines))
    if a.json:
        try:
            old = json.load(open(a.json))
        except (OSError, ValueError):
            old = {}
        old.update(res)
        json.dump(old, open(a.json, 'w'), indent=1, sort_keys=True)


def cmd_counters(a):
    acc = defaultdict(lambda: defaultdict(list))  # kernel -> counter -> per-dispatch values
    for path in a.csv:
        per = defaultdict(float)  # (dispatch, kernel, counter) -> value summed over dimensions
        for r in csv.DictReader(open(path)):
            key = (r.get('Dispatch_Id', r.get('Correlation_Id')), short_name(r['Kernel_Name']),
                   r['Counter_Name'])
            per[key] += float(r['Counter_Value'])
        for (_, k, c), v in per.items():
            acc[k][c].append(v)
    want = a.kernels.split(',') if a.kernels else None
    out = []
    if a.title:
        out.append('# %s\n' % a.title)
    if a.cmd:
        out.append('Commands:\n\n```\n%s\n```\n' % a.cmd)
    for k in sorted(acc):
        if not k.startswith('k_') or (want and k.split('<')[0] not in want and k not in want):
            continue
        c = {n: sum(v) / len(v) for n, v in acc[k].items()}
        out.append('## %s (%d dispatches)\n' % (k, max(len(v) for v in acc[k].values())))
        out.append('| counter | mean per dispatch |')
        out.append('|---|---|')
        for n in sorted(c):
            out.append('| %s | %.4g |' % (n, c[n]))
        waves = c.get('SQ_WAVES')
        der = []
        if waves:
            for n in ('SQ_INSTS_VALU', 'SQ_INSTS_MFMA', 'SQ_INSTS_LDS', 'SQ_INSTS_VMEM_RD',
                      'SQ_INSTS_VMEM_WR', 'SQ_INSTS_SALU', 'SQ_INSTS_SMEM'):
                if n in c:
                    der.append(('%s per wave' % n, c[n] / waves))
        busy = c.get('SQ_BUSY_CYCLES')
        wc = c.get('SQ_WAVE_CYCLES')
        if wc:
            for n in ('SQ_ACTIVE_INST_VALU', 'SQ_ACTIVE_INST_LDS', 'SQ_ACTIVE_INST_VMEM',
                      'SQ_ACTIVE_INST_ANY', 'SQ_WAIT_INST_ANY', 'SQ_WAIT_ANY', 'SQ_WAIT_INST_LDS'):
                if n in c:
                    der.append(('%s / SQ_WAVE_CYCLES' % n, c[n] / wc))
        if 'SQ_LDS_BANK_CONFLICT' in c and 'SQ_ACTIVE_INST_LDS' in c and c['SQ_ACTIVE_INST_LDS']:
            der.append(('SQ_LDS_BANK_CONFLICT / SQ_ACTIVE_INST_LDS',
                        c['SQ_LDS_BANK_CONFLICT'] / c['SQ_ACTIVE_INST_LDS']))
        if busy and 'SQ_VALU_MFMA_BUSY_CYCLES' in c:
            der.append(('SQ_VALU_MFMA_BUSY_CYCLES / SQ_BUSY_CYCLES',
                        c['SQ_VALU_MFMA_BUSY_CYCLES'] / busy))
        if der:
            out.append('\n| derived | value |')
            out.append('|---|---|')
            for n, v in der:
                out.append('| %s | %.4g |' % (n, v))
        out.append('')
    print('\n'.join(out))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest='cmd_name', required=True)
    s = sub.add_parser('stats')
    s.add_argument('csv')
    s.add_argument('--title', default='')
    s.add_argument('--cmd', default='')
    t = sub.add_parser('traffic')
    t.add_argument('fetch')
    t.add_argument('write')
    t.add_argument('--copy-bytes', type=float, required=True)
    t.add_argument('--key-suffix', default='')
    t.add_argument('--json', default='')
    t.add_argument('--title', default='')
    t.add_argument('--cmd', default='')
    w = sub.add_parser('window')
    w.add_argument('csv')
    w.add_argument('--kernel', required=True)
    w.add_argument('--skip', type=int, default=0)
    w.add_argument('--count', type=int, required=True)
    c = sub.add_parser('counters')
    c.add_argument('csv', nargs='+')
    c.add_argument('--kernels', default='')
    c.add_argument('--title', default='')
    c.add_argument('--cmd', default='')
    a = ap.parse_args(argv)
    if a.cmd_name == 'stats':
        cmd_stats(a)
    elif a.cmd_name == 'window':
        cmd_window(a)
    elif a.cmd_name == 'counters':
        cmd_counters(a)
    else:
        cmd_traffic(a)


if __name__ == '__main__':
    main()
