"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):
the CPU oracle (pair test, energies, the serial NVT / NPT sampler, the
generator, .rrr files) and the C host programs' parts (input reader, .rrr
reader/writer, output files, the bond-list writer), each driven through its
normal paths; any sanitizer report fails the test."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
HOST = os.path.join(ROOT, "patchyparticles_amd", "csrc", "host")
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
       "-ffp-contract=off", "-std=gnu99"]

ORACLE_DRIVER = r"""
#include <stdio.h>
#include <stdlib.h>
#include "kf_oracle.h"
int main(void) {
	int n = 400;
	double *r = malloc(sizeof(double) * 3 * n), *O = malloc(sizeof(double) * 9 * n), box[3];
	int32_t *b1 = malloc(sizeof(int32_t) * n), *b2 = malloc(sizeof(int32_t) * n);
	uint8_t *o1 = malloc(n), *o2 = malloc(n);
	if(orc_generate(n, 0.4, 7, r, O, box) != 0) return 2;
	long long e1 = orc_energy_cells(n, r, O, box, 0.12, 0.92, b1, o1);
	long long e2 = orc_energy_brute(n, r, O, box, 0.12, 0.92, b2, o2);
	if(e1 != e2) return 3;
	for(int ens = 0; ens <= 5; ens += 5) {
		orc_params p = {ens, 0.15, 0.05, 0.1, 0.12, 0.92, 0.3, 0.05, 11};
		orc_sim *s = orc_sim_create(&p, n, r, O, box);
		if(!s || orc_sim_init_energy(s)) return 4;
		orc_sim_sweeps(s, 30);
		if(orc_sim_check_energy(s) != 0 || orc_sim_cells_count(s) != n) return 5;
		orc_sim_get(s, r, O, box);
		orc_sim_destroy(s);
	}
	if(orc_rrr_save("san.rrr", 30, n, box, r, O) != 0) return 6;
	long long step = 0;
	double *rr = malloc(sizeof(double) * 3 * n), *OO = malloc(sizeof(double) * 9 * n), bb[3];
	if(orc_rrr_load("san.rrr", &step, bb, rr, OO, n) != n || step != 30) return 7;
	free(rr);
	free(OO);
	printf("ok %lld\n", e1);
	free(r); free(O); free(b1); free(b2); free(o1); free(o2);
	return 0;
}
"""

HOST_DRIVER = r"""
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "pp_input.h"
#include "pp_rrr.h"
#include "pp_output.h"
int main(int argc, char **argv) {
	ppi_file *in = ppi_load(argv[1]);
	if(!in) return 2;
	char s[256];
	int v = 0;
	double x = 0.;
	long long w = 0;
	ppi_string(in, "Name", s, sizeof s, 0);
	ppi_int(in, "Dup", &v, 0);
	ppi_double(in, "Temperature", &x, 0);
	ppi_llint(in, "Steps", &w, 0);
	ppi_int(in, "Missing", &v, 0);
	ppo o;
	ppo_init(&o, in, 0, argv[2]);
	pprrr c;
	char err[512];
	if(pprrr_read(argv[2], &c, err, sizeof err) != 0) return 3;
	if(pprrr_write("copy.rrr", 5, c.n, c.box, c.r, c.O) != 0) return 4;
	int *cnt = calloc(c.n, sizeof(int)), *part = calloc((size_t) c.n * 4, sizeof(int));
	for(int i = 0; i + 1 < c.n; i += 2) {
		cnt[i] = cnt[i + 1] = 1;
		part[4 * i] = i + 1;
		part[4 * (i + 1)] = i;
	}
	const long long gc_acc[2] = {0, 0}, gc_try[2] = {1, 1};
	ppo_print(&o, 0, c.n, -1.0, 100.0, c.box, 3, 4, 0, 1, gc_acc, gc_try);
	ppo_print_bonds(&o, 0, c.n, c.box, c.r, 0.12, cnt, part, 4);
	free(cnt);
	free(part);
	pprrr_free(&c);
	ppo_free(&o);
	ppi_free(in);
	printf("ok\n");
	return 0;
}
"""


def _build(tmp_path, name, src, files, incs):
    c = tmp_path / f"{name}.c"
    c.write_text(src)
    exe = tmp_path / name
    subprocess.run(["gcc", *SAN, *[f"-I{i}" for i in incs], str(c), *files, "-lm", "-o", str(exe)], check=True)
    return str(exe)


def _run(exe, args, cwd):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    res = subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True, env=env, timeout=300)
    report = "ERROR: AddressSanitizer" in res.stderr or "runtime error" in res.stderr or "LeakSanitizer" in res.stderr
    assert res.returncode == 0 and not report, res.stderr[-3000:]
    return res.stdout


def test_oracle_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "orc_san", ORACLE_DRIVER, [os.path.join(ORACLE, "kf_oracle.c")], [ORACLE])
    assert _run(exe, [], tmp_path).startswith("ok")


def test_host_parts_under_sanitizers(tmp_path, oracle):
    exe = _build(tmp_path, "host_san", HOST_DRIVER,
                 [os.path.join(HOST, f) for f in ("pp_input.c", "pp_rrr.c", "pp_output.c")], [HOST])
    r, O, box = oracle.generate(64, 0.2, seed=3)
    oracle.rrr_save(str(tmp_path / "start.rrr"), 0, box, r, O)
    (tmp_path / "bonds").mkdir()
    (tmp_path / "input").write_text(
        "Temperature = 0.1\nSteps = 10\nDup = 1\nDup = 2\nName = a b\nno equals\n"
        "Print_every = 1\nSave_every = 1\nConfiguration_folder = .\nPrint_bonds = 1\nBonds_folder = bonds\n")
    assert _run(exe, ["input", "start.rrr"], tmp_path).strip().endswith("ok")
    assert (tmp_path / "bonds" / "bonds_0.dat").exists() and (tmp_path / "copy.rrr").exists()
