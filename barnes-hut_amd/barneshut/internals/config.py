"""INI configuration singleton, python-bh/barneshut/internals/config.py:8-29
(same static get/getint interface and the same default file name)."""
from configparser import ConfigParser
from pathlib import Path

here = Path(__file__).parent
configs_dir = here / ".." / ".." / "configs"


class Config:

    def __init__(self, ini_path=None):
        default_ini = configs_dir / "default.ini"
        self.parser = ConfigParser()
        read = self.parser.read(ini_path if ini_path is not None else default_ini)
        if ini_path is not None and not read:
            raise FileNotFoundError(f"config file {ini_path} not found")

    @staticmethod
    def read_file(ini_path):
        global config_instance
        config_instance = Config(ini_path)

    @staticmethod
    def get(*args, **kwargs):
        return config_instance.parser.get(*args, **kwargs)

    @staticmethod
    def getint(*args, **kwargs):
        return config_instance.parser.getint(*args, **kwargs)

    @staticmethod
    def getfloat(*args, **kwargs):
        return config_instance.parser.getfloat(*args, **kwargs)


config_instance = None
