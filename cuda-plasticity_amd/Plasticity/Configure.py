"""Run configuration helpers (Plasticity/Configure.py): a dict with
attribute access and its JSON form (.conf files in run archives)."""
import json


class Configuration(dict):
    """The reference also stamps the git hash of its own checkout into the
    configuration; there is none to stamp here, so "hash" is left as given."""

    def __getattr__(self, name):
        try:
            return self[name]
        except KeyError:
            raise AttributeError(name)

    def __setattr__(self, name, value):
        self[name] = value


def open_if_not(filename, mode):
    if not isinstance(filename, str) or hasattr(filename, "read"):
        return filename
    return open(filename, mode)


def write_json(dct, filename):
    with open(filename, "w") as fi:
        fi.write(json.dumps(dct))


def read_json(jsonfile):
    f = open_if_not(jsonfile, 'r')
    text = f.read()
    if isinstance(text, bytes):
        text = text.decode()
    return json.loads(text)
