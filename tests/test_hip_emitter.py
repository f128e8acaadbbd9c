"""Text-level checks of the HIP emitter's fused kernels (CPU, no hipcc).

The GPU parity tests prove the numbers; these pin the layout decisions the
numbers depend on, so that a change which silently drops one shows up here
before a GPU run: per-lane LDS rows are used at both sizes (G rows when
every species fits, concentration rows otherwise), each reaction block's
LDS reads are issued ahead of it, the fused FP64 bodies call no library
transcendental, and the LDS constant-table pass keeps every value.
"""

import re

import pytest

from create_rate_subs_amd import hip_emitter, workloads
from create_rate_subs_amd.create_rate_subs import load_mechanism

_BODIES = ('wdot_f64_state', 'wdot_mixed_state', 'dydt_f64_state')


def _ir(name):
    mpath, tpath = workloads.mechanism_files(name)
    _e, specs, reacs, _fwd = load_mechanism(mpath, tpath)
    return hip_emitter.build_ir(specs, reacs, name=name)


def _bodies(name, opts=None):
    """Source and per-state bodies, with the fused dydt kernel emitted
    (option dydt='fused') so that its body is checked beside the wdot
    ones."""
    text = hip_emitter.emit_hip_source(_ir(name),
                                       dict(opts or {}, dydt='fused'))
    out = {}
    for fn in _BODIES:
        i = text.index('void ' + fn)
        out[fn] = text[i:text.index('\n}\n', i)]
    return text, out


@pytest.mark.parametrize('name,rows,ref,table,waves', [
    ('mech21pc', 'sGl', 'gl', True, None),
    ('mech53', 'sGl', 'gl', False, 2),
    ('mech111', 'sCl', 'cl', False, None)])
def test_lane_lds_rows_and_prefetch(name, rows, ref, table, waves):
    text, bodies = _bodies(name)
    # 40-56 species: two waves per SIMD with 39 G rows per wave
    bound = '__launch_bounds__(kWdotBlock{}) void k_wdot_f64('.format(
        ', {}'.format(waves) if waves else '')
    assert bound in text
    assert 'constexpr int kWdotBlock = 256;' in text
    if waves:
        assert '__shared__ double sGl[kWdotBlock / 64][39][64];' in text
        assert '__launch_bounds__(crs::kBlock) void k_dydt_f64(' in text
    for fn, b in bodies.items():
        assert '__shared__ double {}['.format(rows) in b, fn
        # every block's reads come from a prefetch group outside the block
        blocks = re.findall(r'\{  // rxn \d+\n(.*?)\n    \}', b, re.S)
        assert blocks, fn
        assert not any(re.search(r'\b{}\['.format(ref), blk) for blk in blocks)
        assert re.search(r'const double {}\d+_\d+ = {}p\d+\['.format(ref, ref), b)
        # the NASA table rides along where the H/G rows leave room for it
        # (the dydt kernel stays at one wave per SIMD, where they do)
        want = table or (waves is not None and fn == 'dydt_f64_state')
        assert ('crs_nasa_lds' in b) == want, fn


def test_two_wave_layout_keeps_late_retiring_species_in_lds():
    """With 39 of 53 G rows in LDS, the 14 G_k left in registers are those
    of the species whose last reaction comes first in the emitted order."""
    ir = _ir('mech53')
    _text, bodies = _bodies('mech53')
    b = bodies['wdot_f64_state']
    stored = {int(k) for k in re.findall(r'gb\[\d+ \* 64\] = G(\d+);', b)}
    assert len(stored) == 39
    rxns = hip_emitter._retiring_order(ir)
    end = {k: p for p, x in enumerate(rxns) for k, _nu in x.spec}
    in_regs = set(range(ir.nsp)) - stored
    assert max(end.get(k, -1) for k in in_regs) <= min(end[k] for k in stored)


@pytest.mark.parametrize('name', ['mech53', 'mech111', 'mech21pc', 'mech17tm',
                                  'mech19fx'])
def test_fused_fp64_bodies_call_no_library_transcendental(name):
    _text, bodies = _bodies(name)
    for fn in ('wdot_f64_state', 'dydt_f64_state'):
        calls = re.findall(r'(?<![\w:])(exp|exp10|log|log10|pow)\(',
                           bodies[fn])
        assert calls == [], (fn, sorted(set(calls)))


def test_prefetch_moves_every_read_ahead():
    lines = ['double w[3];']
    for j in range(5):
        lines += ['{  // rxn %d' % j,
                  '    CRS_LDS double *const gl = crs::opaque_lds(gb);',
                  '    const double q = gl[{} * 64] * gl[{} * 64];'.format(
                      j, (j + 2) % 5),
                  '    w[0] = q;', '}', 'W[i] = w[0];']
    for dist in (0, 1, 3, 7):
        out = hip_emitter._prefetch_lane_lds(lines, [('gb', 'gl', True)],
                                             dist)
        defined = {}
        for p, l in enumerate(out):
            mo = re.match(r'const double (gl\d+_\d+) = glp(\d+)\[(\d+) \* 64\];',
                          l)
            if mo:
                defined[mo.group(1)] = (p, int(mo.group(2)), int(mo.group(3)))
        starts = [p for p, l in enumerate(out) if l.startswith('{  // rxn')]
        for j, s in enumerate(starts):
            used = set(re.findall(r'gl\d+_\d+', out[s + 1]))
            assert used == {'gl{}_{}'.format(j, j),
                            'gl{}_{}'.format(j, (j + 2) % 5)}
            for v in used:
                p, blk, _k = defined[v]
                assert blk == j and p < s
                # issued before the block dist places earlier (or the top)
                if j - dist >= 0:
                    assert p < starts[j - dist]
        assert sum(l.startswith('W[i]') for l in out) == 5


def test_constant_table_keeps_every_value():
    lines = ['{  // rxn 0',
             '    const double x = 28.279682783733417 + -0.196 * log2T - '
             '19939.52130381234 * iT;',
             '    const double y = 2.0 * x + 0.5 - 1.0e-300 + 0.196;',
             '}',
             'const double z = 3.25;']
    table = []
    out = hip_emitter._block_consts_to_lds(lines, table)
    assert out[1] == '    CRS_LDS double *const kt = crs::opaque_lds(ktb);'
    body = '\n'.join(out)
    # inline constants stay; 0.196 shares its slot; outside a block unchanged
    assert '2.0 * x + 0.5' in body and out[-1] == 'const double z = 3.25;'
    assert table == [28.279682783733417, 0.196, 19939.52130381234, 1e-300]

    def back(mo):
        return repr(table[int(mo.group(1))])
    restored = [re.sub(r'kt\[(\d+)\]', back, l) for l in out[2:4]]
    num = r'\d+\.\d+(?:e[+-]?\d+)?|\d+e[+-]?\d+'
    nums = [float(v) for l in restored for v in re.findall(num, l)]
    orig = [float(v) for l in lines[1:3] for v in re.findall(num, l)]
    assert nums == orig


@pytest.mark.parametrize('nsp,two', [(13, False), (39, False), (40, True),
                                     (53, True), (56, True), (57, False),
                                     (111, False)])
def test_two_wave_choice(nsp, two):
    """waves_per_simd='auto': two waves with 39 G rows from 40 to 56
    species, one wave (all rows, or concentration rows) otherwise."""
    o = hip_emitter.resolve_opts(nsp)
    assert (o['waves_per_simd'] == 2) == two
    if two:
        assert o['lds_species'] == 39 and o['kc'] == 'species_hg'
        assert o['nasa'] is None and o['g_reload'] and o['lds_prefetch'] == 1
    else:
        assert o['waves_per_simd'] is None
    # an explicit row count or wave count is left alone
    assert hip_emitter.resolve_opts(nsp, {'lds_species': 30})[
        'waves_per_simd'] is None
    assert hip_emitter.resolve_opts(nsp, {'waves_per_simd': None})[
        'waves_per_simd'] is None
    # so is a layout the two-wave rows do not fit beside
    for opts in ({'nasa': 'lds'}, {'kc': 'species'}):
        assert hip_emitter.resolve_opts(nsp, opts)['waves_per_simd'] is None


@pytest.mark.parametrize('opts', [{'g_store': 'lds'}, {'g_store': 'lds_sync'},
                                  {'c_store': 'lds'}])
def test_block_wide_lds_layouts_are_refused(opts):
    """Block-wide LDS arrays would need a barrier after the per-state early
    exits (partial blocks, the subset kernel): refused."""
    with pytest.raises(ValueError):
        hip_emitter.resolve_opts(53, opts)


@pytest.mark.parametrize('nsp', [13, 53, 111])
@pytest.mark.parametrize('block', [64, 128, 256, 512, 1024])
def test_lds_rows_fit_the_block(nsp, block):
    """The per-lane rows (and the NASA table) of every block resident at
    the chosen occupancy fit the CU's 160 KB, whatever the block size."""
    o = hip_emitter.resolve_opts(nsp, {'block': block})
    wps = o['waves_per_simd'] or 1
    wpb = block // 64
    blocks = max(1, max(4 * wps, wpb) // wpb)
    table = 128 * nsp if o['nasa'] == 'lds' else 0
    used = blocks * (wpb * o['lds_species'] * 512 + table)
    assert used <= 160 * 1024, (o, used)
    if block == 256:
        # the measured defaults
        assert o['lds_species'] == {13: 78, 53: 39, 111: 78}[nsp]


def test_wdot_block_option():
    """The wdot kernels' block size is an emitter option: 128 threads give
    two waves' rows per block and a matching launch; sizes that are not a
    multiple of the wave are refused."""
    ir = _ir('mech13')
    text = hip_emitter.emit_hip_source(ir, {'block': 128})
    assert 'constexpr int kWdotBlock = 128;' in text
    assert 'sGl[kWdotBlock / 64]' in text
    assert 'k_wdot_f64<<<crs::grid_for(c, kWdotBlock), kWdotBlock' in text
    with pytest.raises(ValueError):
        hip_emitter.emit_hip_source(ir, {'block': 100})


def test_row_offsets_rewrite():
    """Fused kernels address [rows][ld] arrays through a uniform row pointer
    and one 32-bit lane offset; each late store forms its own row address."""
    lines = ['const double T = T_[i];', 'y[3] = Y[3 * ld + i];',
             'W[7 * ld + i] = w[7];', '    W[2 * ld + i] = 0.0;']
    out = hip_emitter._row_offsets(lines)
    assert out[0] == 'const uint32_t io = (uint32_t)i * 8u;'
    assert out[1] == 'const double T = crs::ld_row(T_, io);'
    assert out[2] == 'y[3] = crs::ld_row(Y + 3 * ld, io);'
    assert out[3] == 'crs::st_row(W + 7 * crs::fresh_s<1>(ld), io, w[7]);'
    assert out[4] == '    crs::st_row(W + 2 * crs::fresh_s<2>(ld), io, 0.0);'


def test_split_blocks_groups_and_hoists():
    """Every `per` reaction blocks form one `if (crs::go<N>())` group; a
    value a group declares for a later group is declared before it."""
    lines = ['const double a = 1.0;',
             '{  // rxn 0', '    const double kf = crs::exp2_64(x);', '}',
             'CRS_LDS double *const glp1 = crs::opaque_lds(gb);',
             'const double gl1_4 = glp1[4 * 64];',
             '{  // rxn 1', '    const double kf = gl0_1;', '}',
             '{  // rxn 2', '    const double kf = gl1_4;', '}',
             'W[0] = 0.0;']
    out = hip_emitter._split_blocks(lines, 2)
    assert out[0] == 'const crs::ExpK ek = crs::exp_consts();'
    assert '    const double kf = crs::exp2_64k(ek, x);' in out
    g1 = out.index('if (crs::go<1>()) {')
    g2 = out.index('if (crs::go<2>()) {')
    assert out.index('double gl1_4;') < g1 < out.index('gl1_4 = glp1[4 * 64];')
    assert 'const double gl1_4 = glp1[4 * 64];' not in out
    assert out.index('gl1_4 = glp1[4 * 64];') < g2
    assert out[-1] == 'W[0] = 0.0;' and out[-2] == '}'


def test_fused_kernels_use_the_transforms():
    ir = _ir('mech13')
    text = hip_emitter.emit_hip_source(ir)
    body = text[text.index('void wdot_f64_state'):
                text.index('void k_wdot_f64')]
    assert 'crs::ld_row(Y + 0 * ld, io)' in body
    assert 'if (crs::go<1>()) {' in body
    assert 'crs::exp2_64k(ek, ' in body and 'crs::exp2_64(' not in body
    plain = hip_emitter.emit_hip_source(ir, {'split_blocks': 0,
                                             'row_offsets': False})
    body = plain[plain.index('void wdot_f64_state'):
                 plain.index('void k_wdot_f64')]
    assert 'crs::go<' not in body and 'Y[0 * ld + i]' in body


def test_split_dydt_is_the_default():
    """By default crs_dydt runs the rate kernel that stores the dydt's
    scaled rates into dy and then the energy stage, crs_rk4 the same rate
    kernel and the energy + RK4-stage kernel; the fused dydt kernel is not
    emitted, and the headline kernel's body is not changed by any of it."""
    text = hip_emitter.emit_hip_source(_ir('mech53'))
    assert 'void k_dydt_f64(' not in text
    assert 'void k_rk4_combine(' not in text
    assert 'void k_dydt_energy(' in text and 'void k_rk4_stage(' in text
    dydt = text[text.index('int crs_dydt('):text.index('int crs_rk4(')]
    w = dydt.index('k_wdot_dydt<<<')
    e = dydt.index('k_dydt_energy<<<')
    assert w < e
    assert 'check_disjoint(fn, y, NSP, "y", dy, NSP, "dy"' in dydt
    # CRS_MODE_MIXED takes the mixed kernel's rates, the same energy stage
    assert 'check_mode(fn, mode)' in dydt
    assert dydt.index('k_wdot_mixed_dydt<<<') < dydt.index('k_dydt_energy<<<')
    rk4 = text[text.index('int crs_rk4('):]
    rk4 = rk4[:rk4.index('\n}\n')]
    assert rk4.index('k_wdot_dydt<<<') < rk4.index('k_rk4_stage<1><<<')
    # the rate kernel: every rate store scaled by W_k / rho, species k in
    # row k + 1, the bath gas (53rd) in row 0; the FP64 body otherwise the
    # headline kernel's
    def state(name):
        t = text[text.index('void {}('.format(name)):]
        return t[:t.index('\n}\n')]
    f64, rates = state('wdot_f64_state'), state('wdot_dydt_state')
    stores = re.findall(r'crs::st_row\(W \+ (\d+) \* crs::fresh_s<\d+>\(ld\), '
                        r'io, (.*)\);', rates)
    assert sorted(int(r) for r, _v in stores) == list(range(53))
    for r, v in stores:
        m = re.match(r'\(w\[(\d+)\]\) \* \(.* \* irho_r\)$', v)
        assert v == '0.0' or m, v
        if m:
            assert int(r) == (int(m.group(1)) + 1) % 53
    assert 'irho_r' not in f64
    assert f64.count('crs::exp2_64k(') == rates.count('crs::exp2_64k(')
    # the energy stage is not contracted, so the RK4 kernel's stage
    # derivatives are the bits crs_dydt returns
    body = text[text.index('void dydt_energy_state('):]
    body = body[:body.index('\n}\n')]
    assert '#pragma clang fp contract(off)' in body
    assert 'P_' not in body and 'rho' not in body
    # the scaled rates in chunks, each row loaded before its output; row
    # k + 1 carries species k, row 0 the bath gas, written last
    loaded = set()
    puts = []
    for line in body.split('\n'):
        m = re.search(r'const double s(\d+) = crs::ld_row\(S \+ (\d+) \* ld',
                      line)
        if m:
            k, r = int(m.group(1)), int(m.group(2))
            assert r == (k + 1) % 53
            loaded.add(r)
        m = re.search(r'out\.put\((\d+),', line)
        if m:
            puts.append(int(m.group(1)))
            assert puts[-1] in loaded, line
    assert loaded == set(range(53))
    assert puts[-1] == 0 and sorted(puts) == list(range(53))
    fused = hip_emitter.emit_hip_source(_ir('mech53'), {'dydt': 'fused'})
    assert 'void k_dydt_f64(' in fused and 'void k_rk4_combine(' in fused
    assert 'void k_dydt_energy(' not in fused
    with pytest.raises(ValueError):
        hip_emitter.emit_hip_source(_ir('mech13'), {'dydt': 'other'})
