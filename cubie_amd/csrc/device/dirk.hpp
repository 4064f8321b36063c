// Placeholder: the dirk component is added with the implicit families.
#pragma once
