"""Parameter tree with the semantics the reference relies on (Boost.PropertyTree).

IO/ptree.hpp:27-84 strips // and /* */ comments before Boost's read_json; every leaf is a
string converted on access by get<T>("a.b") (IO/ptree.hpp:86-119). This module restates that
behaviour: PTree.get(path, type) with dotted paths and get_child(path).
"""
import json


def remove_json_comments(text):
    """Strip // line comments and /* */ block comments (IO/ptree.hpp:27-84).

    Like the reference, the scan is per line and does not track string literals.
    """
    out = []
    in_block = False
    for line in text.split("\n"):
        pos = 0
        piece = ""
        while pos < len(line):
            if in_block:
                end = line.find("*/", pos)
                if end < 0:
                    pos = len(line)
                    break
                in_block = False
                pos = end + 2
            single = line.find("//", pos)
            block = line.find("/*", pos)
            if single >= 0 and (block < 0 or single < block):
                piece += line[pos:single]
                pos = len(line)
                break
            if block >= 0:
                piece += line[pos:block]
                in_block = True
                pos = block + 2
                continue
            piece += line[pos:]
            break
        out.append(piece)
    return "\n".join(out)


def _to_str(v):
    if isinstance(v, bool):
        return "1" if v else "0"
    return str(v)


class PTree:
    """Dotted-path view of a JSON object whose leaves are read as strings (Boost ptree)."""

    def __init__(self, data=None):
        self._d = {} if data is None else data

    @classmethod
    def read(cls, path):
        with open(path) as f:
            return cls(json.loads(remove_json_comments(f.read())))

    @classmethod
    def from_string(cls, text):
        return cls(json.loads(remove_json_comments(text)))

    def _node(self, path):
        node = self._d
        if path:
            for key in path.split("."):
                if not isinstance(node, dict) or key not in node:
                    raise KeyError(f"No such node ({path})")
                node = node[key]
        return node

    def get(self, path, typ=str, default=None):
        try:
            node = self._node(path)
        except KeyError:
            if default is not None:
                return default
            raise
        if isinstance(node, dict):
            raise KeyError(f"{path} is not a leaf")
        s = _to_str(node).strip()
        if typ is bool:
            return s not in ("0", "false", "")
        if typ is int:
            return int(float(s)) if ("e" in s.lower() or "." in s) else int(s)
        return typ(s)

    def get_child(self, path):
        node = self._node(path)
        if not isinstance(node, dict):
            raise KeyError(f"{path} is not a subtree")
        return PTree(node)

    def put(self, path, value):
        keys = path.split(".")
        node = self._d
        for k in keys[:-1]:
            node = node.setdefault(k, {})
        node[keys[-1]] = _to_str(value)

    def has(self, path):
        try:
            self._node(path)
            return True
        except KeyError:
            return False

    def to_dict(self):
        return json.loads(json.dumps(self._d))
