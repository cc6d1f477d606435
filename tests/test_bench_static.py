"""bench.py names only what it defines, imports or the builtins provide (CPU; catches a removed helper
before a GPU run reaches it)."""
import ast
import builtins
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_calls_only_known_names():
    tree = ast.parse(open(os.path.join(ROOT, 'bench.py')).read())
    known = set(dir(builtins))
    for node in ast.walk(tree):
        if isinstance(node, (ast.FunctionDef, ast.ClassDef)):
            known.add(node.name)
        elif isinstance(node, (ast.Import, ast.ImportFrom)):
            known.update((a.asname or a.name).split('.')[0] for a in node.names)
        elif isinstance(node, ast.Name) and isinstance(node.ctx, ast.Store):
            known.add(node.id)
        elif isinstance(node, ast.arg):
            known.add(node.arg)
    called = {n.func.id for n in ast.walk(tree) if isinstance(n, ast.Call) and isinstance(n.func, ast.Name)}
    assert not sorted(called - known)


def test_host_info():
    import sys
    sys.path.insert(0, ROOT)
    import bench
    info = bench.host_info()
    assert info['os_cpu_count'] >= 1 and isinstance(info['cpu_model'], str)
