'''Writes tests/golden/cart2sph_table.json from orbkit's own cart2sph table.

The table is read as text: orbkit/tools.py is parsed with `ast` and the
literal assigned to `cart2sph` is evaluated by a small evaluator that accepts
only numbers, tuples, lists, unary minus, + - * / and calls of sqrt (nothing
from the file is imported or executed).  The fixture holds, per (l, m), the
Cartesian exponents, their coefficients and the global factor in orbkit's
order (cart2sph[l][l + m]), signs included.

    python tests/golden/make_cart2sph_fixture.py [/root/reference/orbkit/tools.py]
'''
import ast
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = sys.argv[1] if len(sys.argv) > 1 else '/root/reference/orbkit/tools.py'


def value(node):
  if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
    return node.value
  if isinstance(node, ast.Tuple):
    return tuple(value(e) for e in node.elts)
  if isinstance(node, ast.List):
    return [value(e) for e in node.elts]
  if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
    return -value(node.operand)
  if isinstance(node, ast.BinOp):
    a, b = value(node.left), value(node.right)
    ops = {ast.Add: lambda: a + b, ast.Sub: lambda: a - b, ast.Mult: lambda: a * b, ast.Div: lambda: a / b}
    return ops[type(node.op)]()
  if (isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id == 'sqrt'
      and len(node.args) == 1 and not node.keywords):
    return math.sqrt(value(node.args[0]))
  raise ValueError('unexpected node %s at line %d' % (type(node).__name__, node.lineno))


def main():
  tree = ast.parse(open(SRC).read())
  table = None
  for node in tree.body:
    if (isinstance(node, ast.Assign) and len(node.targets) == 1 and isinstance(node.targets[0], ast.Name)
        and node.targets[0].id == 'cart2sph'):
      table = value(node.value)
  if table is None:
    raise SystemExit('no cart2sph literal in %s' % SRC)
  rows = []
  for l, shell in enumerate(table):
    for i, (exps, coeffs, glob) in enumerate(shell):
      rows.append({'l': l, 'm': i - l, 'exps': [list(e) for e in exps], 'coeffs': list(coeffs), 'global': glob})
  out = os.path.join(HERE, 'cart2sph_table.json')
  with open(out, 'w') as f:
    json.dump({'source': 'orbkit/tools.py cart2sph (cart2sph[l][l + m])', 'rows': rows}, f, indent=1)
  print('wrote %s: %d rows' % (out, len(rows)))


if __name__ == '__main__':
  main()
