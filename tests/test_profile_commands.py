"""Every rocprofv3 summary under profiles/ quotes the command that produced
it.  The options those commands pass to bench.py or to
imresh_amd.kernel_profile must still be options the scripts accept, so each
summary can be re-collected from the committed tree; and no command may
profile graph replays (bench legs default to cycle graphs unless run with
--cycle-graphs off; kernel_profile launches eagerly)."""
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parents[1]
SCRIPTS = {
    "bench.py": ROOT / "bench.py",
    "imresh_amd.kernel_profile": ROOT / "imresh_amd" / "kernel_profile.py",
}


def _options(path: pathlib.Path) -> set:
    return set(re.findall(r'add_argument\(\s*"(--[a-z0-9-]+)"', path.read_text()))


def _commands():
    for md in sorted((ROOT / "profiles").glob("*.md")):
        for line in md.read_text().splitlines():
            if "rocprofv3" not in line or " -- " not in line:
                continue
            for script in SCRIPTS:
                i = line.find(script)
                if i >= 0:
                    yield md.name, script, line[i + len(script):]


def test_profile_commands_use_options_the_scripts_accept():
    known = {s: _options(p) for s, p in SCRIPTS.items()}
    assert "--cycle-graphs" in known["bench.py"] and "--size" in known["imresh_amd.kernel_profile"]
    seen = 0
    for name, script, args in _commands():
        used = re.findall(r"(?<![\w-])(--[a-z0-9-]+)", args)
        missing = [u for u in used if u not in known[script]]
        assert not missing, (name, script, missing)
        seen += 1
    assert seen >= 8


def test_profiled_bench_runs_launch_eagerly():
    for name, script, args in _commands():
        if script == "bench.py":
            assert re.search(r"--cycle-graphs\s+off\b", args), name
