"""Call recorder used to run lap functions: every attribute access returns a
callable that forwards (name, *args, **kwargs) to an action, bracketed by
optional pre/post hooks.  Same behaviour as runko/method_wrapper.py:7-89
(argument binding is checked against each hook before anything runs)."""

import inspect


class MethodWrapper:
    def __init__(self, action, pre=None, post=None):
        self._action = action
        self._pre = pre
        self._post = post

    def __getattr__(self, name):
        if name.startswith("__"):
            raise AttributeError(
                f"MethodWrapper invalid method name: {name}\n"
                "In order to prevent unindented effects, attributes starting with `__`"
                "can not be customized.")

        def call(*args, **kwargs):
            for role, fn in (("pre", self._pre), ("action", self._action), ("post", self._post)):
                if fn is None:
                    continue
                try:
                    inspect.signature(fn).bind(name, *args, **kwargs)
                except TypeError:
                    raise TypeError(f"In wrapped `{name}` method, {role} can not handle arguments: "
                                    f"{name}, *{args}, **{kwargs}")
            if self._pre:
                self._pre(name, *args, **kwargs)
            self._action(name, *args, **kwargs)
            if self._post:
                self._post(name, *args, **kwargs)

        return call
