"""Configuration object with runko's semantics (runko/configuration.py:4-56):
ini sections read through ast.literal_eval into attributes; a missing
attribute reads as None, dunder names raise AttributeError."""

import ast
import configparser


class Configuration:
    def __init__(self, config_path):
        self._config_path = config_path
        self.verbose = config_path is not None
        if not config_path:
            return
        parser = configparser.ConfigParser(inline_comment_prefixes=("#", ";"))
        parser.optionxform = str
        if not parser.read(config_path):
            raise ValueError(f"No config found: {config_path}")
        for section in parser.sections():
            for key, value in parser.items(section):
                self.__dict__[key] = ast.literal_eval(value)

    def __getattr__(self, name):
        if name.startswith("__"):
            raise AttributeError(f"Configuration does not define: {name}")
        return None
