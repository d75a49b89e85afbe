"""Exceptions (p4/p4exceptions.py)."""


class P4Error(Exception):
    """Raised for bad input or state; the message is a string or a list of
    lines, as p4's P4Error takes."""

    def __init__(self, msg=""):
        if isinstance(msg, (list, tuple)):
            msg = "\n".join(str(m) for m in msg)
        super().__init__(msg)
        self.msg = msg
