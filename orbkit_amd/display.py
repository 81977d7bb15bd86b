'''Progress messages (orbkit/display.py:28-43), printed unless ``quiet``.'''
import sys

quiet = True


def display(text):
  if not quiet:
    sys.stdout.write(str(text) + '\n')
