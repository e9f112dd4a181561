"""The reference's constant tables, read once from the committed fixture."""
import functools
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


@functools.lru_cache(maxsize=None)
def tables() -> dict:
    with open(os.path.join(HERE, "golden", "reference_tables.json")) as f:
        return json.load(f)
